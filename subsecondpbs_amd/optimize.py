"""Spot-weight optimisation loop on the differentiation case (config-4 objective).

Runs projected Adam steps on the spot weights of one beam: every step is one forward dose
(``pbs_engine_beam_dose``) and one adjoint (``pbs_engine_beam_adjoint``) through
``SpotWeightDose``; weights are clamped at zero after each update. Prints one JSON line with
the objective trajectory and the time per step.

    python -m subsecondpbs_amd.optimize --steps 50
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from . import fixtures as F
from .autograd import beam_dose, objective
from .engine import Engine


def setup(case: dict):
    eng = Engine(case["energy"])
    eng.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    eng.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    beam = dict(case["beams"][0])
    w0 = torch.tensor(np.asarray(beam["spot_weights"], np.float32))
    with torch.no_grad():
        d = beam_dose(w0, eng, beam)
    t = torch.from_numpy(case["target_mask"]).cuda()
    o = torch.from_numpy(case["oar_mask"]).cuda()
    # the dose is linear in the weights: scale them to the prescription's mean target dose
    scale = case["prescription"] / float(d[t].mean().item())
    return eng, beam, w0 * scale, t, o


def run(steps: int = 50, lr_frac: float = 0.05, seed: int = 1, mode: str = "graph") -> dict:
    """mode: "full" (every step runs the whole chain: trace, fill, ... and the full adjoint),
    "resident" (the beam kept resident: weight-only forward and adjoint from the state) or
    "graph" (resident, with the forward replayed as a captured HIP graph)."""
    case = F.make_config4(seed)
    eng, beam, w, t, o = setup(case)
    resident = None
    if mode in ("resident", "graph"):
        nx, ny, nz = case["dose_dims"]
        b0 = dict(beam)
        b0["spot_weights"] = w.numpy().copy()
        tmp = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
        resident = eng.keep_beam(b0, tmp)
        eng.sync()
        resident.use_graph(mode == "graph")
    w = w.clone().requires_grad_(True)
    opt = torch.optim.Adam([w], lr=lr_frac * float(w.detach().mean()))
    pres, of = case["prescription"], case["oar_factor"]
    hist, times = [], []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        opt.zero_grad()
        J = objective(beam_dose(w, eng, beam, resident), t, o, pres, of)
        J.backward()
        opt.step()
        with torch.no_grad():
            w.clamp_(min=0.0)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
        hist.append(float(J.item()))
    with torch.no_grad():
        Jf = float(objective(beam_dose(w, eng, beam, resident), t, o, pres, of).item())
    if resident is not None:
        resident.close()
    eng.close()
    return {"case": case["name"], "steps": steps, "spots": int(w.numel()), "objective_first": hist[0],
            "objective_last": Jf, "objective": hist, "ms_per_step": float(np.mean(times[1:] if steps > 1 else times)),
            "ms_per_step_p50": float(np.percentile(times, 50)), "total_ms": float(np.sum(times)),
            "mode": mode, "step": "projected Adam; forward dose + hand-written adjoint per step"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--lr-frac", type=float, default=0.05)
    ap.add_argument("--mode", choices=("full", "resident", "graph"), default="graph")
    a = ap.parse_args()
    out = run(a.steps, a.lr_frac, mode=a.mode)
    out["objective"] = [round(v, 6) for v in out["objective"]]
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
