"""Adjoint of the HIP engine with respect to the spot weights (config-4 case: config-1 beam
with a target / OAR objective).

Checks: the GPU gradient against the FP64 oracle's adjoint; bitwise repeatability;
<g, dose(w)> = <w, A^T g> on the
GPU (the dose is linear in the weights for fixed masks); central finite differences of the GPU
forward on 64 seeded spots with h = 1e-3 w; torch.autograd through SpotWeightDose; a few
descent steps lower the objective.
"""
import numpy as np
import pytest

import oracle as O
from subsecondpbs_amd import fixtures as F

torch = pytest.importorskip("torch")
gpu = pytest.mark.gpu


def _engine(case):
    from subsecondpbs_amd.engine import Engine

    eng = Engine(case["energy"])
    eng.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    eng.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    return eng


def _dose(eng, case, beam):
    nx, ny, nz = case["dose_dims"]
    d = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    eng.beam_dose(beam, d)
    eng.sync()
    return d


@pytest.fixture(scope="module")
def cfg4():
    case = F.make_config4()
    beam = dict(case["beams"][0])
    eng = _engine(case)
    d = _dose(eng, case, beam).cpu().numpy().astype(np.float64)
    scale = case["prescription"] / d[case["target_mask"]].mean()
    beam["spot_weights"] = (np.asarray(beam["spot_weights"], np.float64) * scale).astype(np.float32)
    case["beams"] = [beam]
    dose = _dose(eng, case, beam)
    t = torch.from_numpy(case["target_mask"]).cuda()
    o = torch.from_numpy(case["oar_mask"]).cuda()
    g = (2.0 * (dose - case["prescription"]) * t + 2.0 * case["oar_factor"] * dose * o).float().contiguous()
    print(f"\n[gpu-test] config4 weights scaled by {scale:.4e}; mean target dose "
          f"{dose[t].mean().item():.4f}", flush=True)
    return case, beam, eng, dose, g


@gpu
def test_adjoint_matches_fp64_oracle(cfg4):
    case, beam, eng, dose, g = cfg4
    gw = eng.beam_adjoint(beam, g).astype(np.float64)
    ref = O.beam_adjoint(case, g.cpu().numpy().astype(np.float64))
    err = np.abs(gw - ref).max() / np.abs(ref).max()
    print(f"adjoint vs FP64 oracle: max|dg|/max|g| = {err:.3e}")
    assert err < 2e-3


@gpu
def test_adjoint_bitwise_repeatable(cfg4):
    """The transposed transfer sums in 64-bit fixed point, so the order of its atomics does
    not reach the result: repeated adjoints are bitwise equal. A scaled gradient (another
    fixed-point scale) is repeatable too, and scales the result."""
    case, beam, eng, dose, g = cfg4
    a = eng.beam_adjoint(beam, g)
    for _ in range(3):
        assert np.array_equal(eng.beam_adjoint(beam, g), a)
    g4 = (4.0 * g).contiguous()
    b = eng.beam_adjoint(beam, g4)
    assert np.array_equal(eng.beam_adjoint(beam, g4), b)
    np.testing.assert_allclose(b, 4.0 * a, rtol=1e-6, atol=1e-6 * float(np.abs(a).max()))


@gpu
def test_adjoint_dot_product_on_gpu(cfg4):
    case, beam, eng, dose, g = cfg4
    gw = eng.beam_adjoint(beam, g).astype(np.float64)
    w = np.asarray(beam["spot_weights"], np.float64)
    lhs = float((g.double() * dose.double()).sum().item())
    rhs = float((w * gw).sum())
    print(f"<g, dose> = {lhs:.6e}  <w, grad> = {rhs:.6e}")
    assert abs(lhs - rhs) <= 1e-4 * abs(lhs)


@gpu
def test_adjoint_central_differences_64_spots(cfg4):
    case, beam, eng, dose, g = cfg4
    gw = eng.beam_adjoint(beam, g).astype(np.float64)
    w = np.asarray(beam["spot_weights"], np.float32)
    gd = g.double()
    fd = np.zeros(len(case["fd_spots"]))
    for i, s in enumerate(case["fd_spots"]):
        h = np.float32(1e-3) * w[s]
        wp, wm = w.copy(), w.copy()
        wp[s] += h
        wm[s] -= h
        bp, bm = dict(beam), dict(beam)
        bp["spot_weights"], bm["spot_weights"] = wp, wm
        dp = _dose(eng, case, bp).double()
        dm = _dose(eng, case, bm).double()
        fd[i] = float((gd * (dp - dm)).sum().item()) / (float(wp[s]) - float(wm[s]))
    ref = gw[case["fd_spots"]]
    rel = np.linalg.norm(fd - ref) / np.linalg.norm(ref)
    print(f"central differences on 64 spots: |fd - adjoint| / |adjoint| = {rel:.3e}")
    assert rel < 2e-2


@gpu
def test_autograd_function_and_descent(cfg4):
    from subsecondpbs_amd.autograd import beam_dose, objective

    case, beam, eng, dose, g = cfg4
    t = torch.from_numpy(case["target_mask"]).cuda()
    o = torch.from_numpy(case["oar_mask"]).cuda()
    w = torch.tensor(np.asarray(beam["spot_weights"], np.float32), requires_grad=True)
    d = beam_dose(w, eng, beam)
    J = objective(d, t, o, case["prescription"], case["oar_factor"])
    J.backward()
    direct = eng.beam_adjoint(beam, g).astype(np.float64)
    # the objective's gradient is formed in float64 by autograd, g in float32 here; the BEV
    # gradient uses float atomics: agreement to rounding
    got = w.grad.numpy().astype(np.float64)
    assert np.abs(got - direct).max() <= 1e-5 * np.abs(direct).max()
    # projected gradient descent with a fixed step lowers the objective
    J0 = float(J.item())
    x = w.detach().clone()
    step = 0.5 * float(x.abs().max()) / float(w.grad.abs().max())
    for _ in range(3):
        x.requires_grad_(True)
        Jx = objective(beam_dose(x, eng, beam), t, o, case["prescription"], case["oar_factor"])
        Jx.backward()
        with torch.no_grad():
            x = (x - step * x.grad).clamp_min(0.0)
    Jn = float(objective(beam_dose(x, eng, beam), t, o, case["prescription"], case["oar_factor"]).item())
    print(f"objective {J0:.4e} -> {Jn:.4e} after 3 projected steps")
    assert Jn < J0
