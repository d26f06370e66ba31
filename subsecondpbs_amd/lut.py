"""Physics look-up tables: proton cumulative depth dose, HU->density, HU->stopping
power, density->1/X0 and the nuclear-halo weight / sigma^2 tables.

Restates ``energyReader`` (src/utils/energy_reader.cpp:21-185): whitespace-separated
text files read token by token in the same order and with the same sizes.

* ``proton_cumul_ddd_data.txt``: ``nSamples nEnergies`` then ``nEnergies`` energies
  (MeV/u), ``nEnergies`` Bragg-peak depths (mm), ``nEnergies`` depth scale factors
  (samples per mm) and ``nEnergies*nSamples`` cumulative IDD values, energy-major
  (energy_reader.cpp:50-72).
* ``density_Schneider2000_adj.txt``, ``HU_to_SP_H&N_adj.txt``, ``radiation_length.txt``:
  ``n scale`` then ``n`` values (energy_reader.cpp:74-132).
* ``nuclear_weights_and_sigmas_fit.txt``: header identical to the CIDD file, then the
  weight matrix and the sigma^2 matrix, both energy-major (energy_reader.cpp:134-181).

The returned dict uses the key names of the Python binding's ``energy`` argument
(src/bindings/raytracedicom_pybind.cpp:758-794, 1825-1847).
"""
from __future__ import annotations

import os
from functools import lru_cache

import numpy as np

TABLES_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables")

CIDD_FILE = "proton_cumul_ddd_data.txt"
DENSITY_FILE = "density_Schneider2000_adj.txt"
SP_FILE = "HU_to_SP_H&N_adj.txt"
RRL_FILE = "radiation_length.txt"
RRL_WATER_FILE = "radiation_length_inc_water.txt"
NUC_FILE = "nuclear_weights_and_sigmas_fit.txt"


def _tokens(path: str) -> np.ndarray:
    with open(path, "r") as fh:
        return np.array(fh.read().split(), dtype=np.float64)


def _read_scaled_vector(path: str):
    tok = _tokens(path)
    n = int(tok[0])
    scale = float(tok[1])
    vec = tok[2 : 2 + n].astype(np.float32)
    if vec.size != n:
        raise ValueError(f"{path}: expected {n} samples, found {vec.size}")
    return n, np.float32(scale), vec


@lru_cache(maxsize=4)
def _load_cached(tables_dir: str, water_rrl: bool, nuclear: bool):
    tok = _tokens(os.path.join(tables_dir, CIDD_FILE))
    n_samples = int(tok[0])
    n_energies = int(tok[1])
    off = 2
    energies = tok[off : off + n_energies].astype(np.float32)
    off += n_energies
    peaks = tok[off : off + n_energies].astype(np.float32)
    off += n_energies
    scales = tok[off : off + n_energies].astype(np.float32)
    off += n_energies
    cidd = tok[off : off + n_energies * n_samples].astype(np.float32)
    if cidd.size != n_energies * n_samples:
        raise ValueError("CIDD table truncated")
    cidd = cidd.reshape(n_energies, n_samples)

    nd, ds, dv = _read_scaled_vector(os.path.join(tables_dir, DENSITY_FILE))
    ns, ss, sv = _read_scaled_vector(os.path.join(tables_dir, SP_FILE))
    nr, rs, rv = _read_scaled_vector(
        os.path.join(tables_dir, RRL_WATER_FILE if water_rrl else RRL_FILE))

    out = {
        "n_energies": n_energies,
        "n_energy_samples": n_samples,
        "energies_per_u": energies,
        "peak_depths": peaks,
        "scale_facts": scales,
        "cidd_matrix": cidd,
        "n_density_samples": nd,
        "density_scale_fact": ds,
        "density_vector": dv,
        "n_sp_samples": ns,
        "sp_scale_fact": ss,
        "sp_vector": sv,
        "n_rrl_samples": nr,
        "rrl_scale_fact": rs,
        "rrl_vector": rv,
    }
    if nuclear:
        ntok = _tokens(os.path.join(tables_dir, NUC_FILE))
        if int(ntok[0]) != n_samples or int(ntok[1]) != n_energies:
            raise ValueError("nuclear table axis differs from the CIDD table")
        o = 2
        ne = ntok[o : o + n_energies]
        o += n_energies
        npk = ntok[o : o + n_energies]
        o += n_energies
        nsc = ntok[o : o + n_energies]
        o += n_energies
        # energy_reader.cpp:158-176: axes must agree within 0.01
        if (np.abs(ne - energies) > 0.01).any() or (np.abs(npk - peaks) > 0.01).any() \
                or (np.abs(nsc - scales) > 0.01).any():
            raise ValueError("nuclear table energies/peaks/scales differ from CIDD table")
        w = ntok[o : o + n_energies * n_samples].astype(np.float32)
        o += n_energies * n_samples
        s2 = ntok[o : o + n_energies * n_samples].astype(np.float32)
        out["nuclear_weight_matrix"] = w.reshape(n_energies, n_samples)
        out["nuclear_sq_sigma_matrix"] = s2.reshape(n_energies, n_samples)
    return out


def load_energy_tables(tables_dir: str | None = None, *, water_rrl: bool = False,
                       nuclear: bool = False) -> dict:
    """Return the energy/LUT dict (a fresh shallow copy; arrays are shared and read-only)."""
    d = _load_cached(os.path.abspath(tables_dir or TABLES_DIR), bool(water_rrl), bool(nuclear))
    out = dict(d)
    for v in out.values():
        if isinstance(v, np.ndarray):
            v.setflags(write=False)
    return out
