"""Project-defined carbon-ion (12C) tables for the final-dose engine. NOT reference data.

The reference ships proton tables only (energy_reader.cpp:135-181) and clamps carbon energies
into the proton axis (wrapper:3951-3970); its ``kernelSuperpositionCarbon``
(superposition.cu:88-155) is dead code. Carbon beams here run the ordinary chain (trace,
convolution, fill, tile radii, superposition, transfer, and the halo branch) on a table set
built by this module in the proton tables' layout. The numbers follow a small, stated model,
not measured data, and every result computed with them is flagged as using a project LUT.

Model (E in MeV/u, depths in mm of water):

* range-energy (Bragg-Kleeman): R(E) = ALPHA E^P with ALPHA = 8.0e-3 mm, P = 1.75, which puts
  100 / 290 / 400 MeV/u at 25 / 162 / 286 mm; the fill's residual energy is its inverse,
  E(r) = ALPHA^(-1/P) r^(1/P) (``ion`` constants below);
* stopping along the path: dE/dz proportional to (R - z)^(1/P - 1) (the derivative of the law);
* primary fluence: 1 - KAPPA z / R with KAPPA = 0.45 (R / 290 mm): nuclear loss to fragments,
  about 45 % at 16 cm;
* range straggling: a Gaussian of sigma 0.35 % of R (the proton 1.2 % scaled by 1/sqrt(12));
* fragment tail past the peak: 12 % of the peak dose decaying as exp(-(z - R) / (0.4 R));
* cumulative depth dose (CIDD): 1024 samples with the peak at sample 819 (the proton tables'
  invariant peak x scale = 819), normalised to 1.5e-7 x 12 E at the last sample (the proton
  tables' CIDD(end) / E);
* halo weight: the fraction of the step's energy in fragments, KAPPA z / R before the peak and
  the tail share past it; halo sigma^2 = (2 mm + 0.03 z)^2;
* ion constants of the fill: rest mass 931.494 MeV per nucleon, multiple scattering scaled by
  (Z / A)^2 = 0.25.

The density, stopping-power ratio and radiation-length vectors are the proton set's (material
properties relative to water).
"""
from __future__ import annotations

import numpy as np

from .lut import load_energy_tables

ALPHA_MM = 8.0e-3
P_EXP = 1.75
KAPPA_290 = 0.45
STRAGGLE = 0.0035
TAIL_FRACTION = 0.12
TAIL_LENGTH = 0.4
N_SAMPLES = 1024
PEAK_SAMPLE = 819.0
CIDD_PER_MEV = 1.5e-7
ION = {"name": "carbon-12 (project-defined table, not reference data)", "charge": 6, "mass_number": 12,
       "range_coef": float(ALPHA_MM ** (-1.0 / P_EXP)), "range_exp": float(1.0 / P_EXP), "rest_mass": 931.494,
       "scatter_scale": (6.0 / 12.0) ** 2}


def carbon_range_mm(e_per_u):
    return ALPHA_MM * np.power(np.asarray(e_per_u, np.float64), P_EXP)


def _idd(z, R):
    """Relative depth dose of one energy on depths z (mm)."""
    kappa = KAPPA_290 * R / 290.0
    d = np.clip(R - z, 1e-6 * R, None)
    stop = np.power(d / R, 1.0 / P_EXP - 1.0)          # dE/dz shape before the end of range
    prim = np.where(z < R, (1.0 - kappa * z / R) * stop, 0.0)
    # straggling: Gaussian smoothing on the sample grid
    sig = STRAGGLE * R
    dz = z[1] - z[0]
    half = int(np.ceil(4.0 * sig / dz))
    k = np.exp(-0.5 * (np.arange(-half, half + 1) * dz / sig) ** 2)
    k /= k.sum()
    prim = np.convolve(np.pad(prim, half, mode="edge"), k, mode="valid")
    peak = prim.max()
    tail = np.where(z > R, TAIL_FRACTION * peak * np.exp(-(z - R) / (TAIL_LENGTH * R)), 0.0)
    frag_pre = np.where(z <= R, kappa * z / R * TAIL_FRACTION * peak, 0.0)
    return prim + tail + frag_pre, prim, tail + frag_pre


def build_carbon_tables(e_min: float = 100.0, e_max: float = 400.0, n_energies: int = 61) -> dict:
    """Energy dict in the proton tables' layout (see module docstring), with ``ion`` constants
    and the halo weight / sigma^2 matrices."""
    base = load_energy_tables()
    E = np.linspace(e_min, e_max, n_energies)
    R = carbon_range_mm(E)
    scale = PEAK_SAMPLE / R                              # samples per mm
    cidd = np.zeros((n_energies, N_SAMPLES))
    nucw = np.zeros((n_energies, N_SAMPLES))
    nucs = np.zeros((n_energies, N_SAMPLES))
    for i in range(n_energies):
        z = np.arange(N_SAMPLES) / scale[i]
        total, prim, frag = _idd(z, R[i])
        c = np.cumsum(total) * (z[1] - z[0])
        cidd[i] = c / c[-1] * CIDD_PER_MEV * 12.0 * E[i]
        with np.errstate(invalid="ignore", divide="ignore"):
            w = np.where(total > 0.0, frag / total, 0.0)
        nucw[i] = np.clip(w, 0.0, 1.0)
        nucs[i] = (2.0 + 0.03 * z) ** 2
    out = dict(base)
    out.update({
        "n_energies": n_energies, "n_energy_samples": N_SAMPLES,
        "energies_per_u": E.astype(np.float32), "peak_depths": R.astype(np.float32),
        "scale_facts": scale.astype(np.float32), "cidd_matrix": cidd.astype(np.float32),
        "nuclear_weight_matrix": nucw.astype(np.float32), "nuclear_sq_sigma_matrix": nucs.astype(np.float32),
        "ion": dict(ION),
    })
    return out
