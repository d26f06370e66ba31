"""Synthetic treatment-plan fixtures (CT, beams, spots) for the pencil-beam dose path.

Every random draw comes from a counter-based splitmix64 stream written here, so the
same seed gives the same phantom and spot weights on every machine and numpy version.

The beam dicts use the key names of the reference Python binding
(src/bindings/raytracedicom_pybind.cpp:1712-1821); the energy dict is
:func:`subsecondpbs_amd.lut.load_energy_tables`.

Conventions (mm throughout, so the reference's length heuristic picks lenToMm = 1,
src/core/raytracedicom_wrapper.cu:2728-2729):

* CT and dose volumes are x-fastest float32 ``[nz][ny][nx]``; corners are voxel centres.
* ``spot_positions`` are gantry-frame (x, y) coordinates on the isocentre plane, which is
  how the reference's spot-lattice path consumes them (wrapper:1366-1456, 3044-3102).
"""
from __future__ import annotations

import math

import numpy as np

from .lut import load_energy_tables

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based RNG: draw ``i`` of stream (seed, name) is splitmix64(key + i)."""

    def __init__(self, seed: int, name: str):
        h = 1469598103934665603
        for ch in name.encode():
            h = ((h ^ ch) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
        base = np.array([(seed * 0x9E3779B97F4A7C15 + h) & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)
        self.key = _splitmix64(base)[0]
        self.counter = 0

    def _raw(self, n: int) -> np.ndarray:
        idx = np.arange(self.counter, self.counter + n, dtype=np.uint64)
        self.counter += n
        with np.errstate(over="ignore"):
            return _splitmix64(idx * np.uint64(0x9E3779B97F4A7C15) + self.key)

    def uniform(self, n: int) -> np.ndarray:
        """Doubles in [0, 1) with 53 random bits."""
        return (self._raw(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        """Box-Muller standard normals."""
        u1 = 1.0 - self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def _grid_centres(dims, res, corner):
    xs = corner[0] + res[0] * np.arange(dims[0], dtype=np.float64)
    ys = corner[1] + res[1] * np.arange(dims[1], dtype=np.float64)
    zs = corner[2] + res[2] * np.arange(dims[2], dtype=np.float64)
    return xs, ys, zs


def subspot_lut(num_layers: int, sigma_sub: float | np.ndarray, offsets_mm=(-4.0, -2.0, 0.0, 2.0, 4.0),
                var_weight: float = 12.0) -> np.ndarray:
    """[L][S][5] subspot table (dx, dy, w, sigma_x, sigma_y); weights ~ exp(-|d|^2/(2 var)), normalised."""
    offs = np.asarray(offsets_mm, dtype=np.float64)
    dx, dy = np.meshgrid(offs, offs, indexing="xy")
    dx = dx.ravel()
    dy = dy.ravel()
    w = np.exp(-(dx * dx + dy * dy) / (2.0 * var_weight))
    w /= w.sum()
    sig = np.broadcast_to(np.asarray(sigma_sub, dtype=np.float64), (num_layers,))
    out = np.zeros((num_layers, dx.size, 5), dtype=np.float32)
    out[:, :, 0] = dx
    out[:, :, 1] = dy
    out[:, :, 2] = w
    out[:, :, 3] = sig[:, None]
    out[:, :, 4] = sig[:, None]
    return out


def single_subspot_lut(spot_sigmas: np.ndarray) -> np.ndarray:
    """[L][1][5] table holding one centred subspot per layer with the layer's spot sigma."""
    L = spot_sigmas.shape[0]
    out = np.zeros((L, 1, 5), dtype=np.float32)
    out[:, 0, 2] = 1.0
    out[:, 0, 3] = spot_sigmas[:, 0]
    out[:, 0, 4] = spot_sigmas[:, 1]
    return out


def beam_basis(gantry_deg: float, couch_deg: float = 0.0):
    """Beam direction and lateral axes for a coplanar gantry angle (IEC-like, about world z).

    Gantry 0 points the beam along +z? No: gantry rotates in the x-y plane here. Gantry 90
    gives a beam along +x with bmX = +y and bmY = +z, so that bmX x bmY = beam direction.
    """
    a = math.radians(gantry_deg)
    bdir = np.array([math.sin(a), -math.cos(a), 0.0])   # 90 deg -> +x ; 0 deg -> -y
    bmy = np.array([0.0, 0.0, 1.0])
    bmx = np.cross(bmy, bdir)  # so that cross(bmx, bmy) = bdir
    bmx /= np.linalg.norm(bmx)
    return bdir, bmx, bmy


def make_beam(energies, spot_xy_per_layer, weights_per_layer, spot_sigmas, *, beam_dir, bm_x, bm_y,
              sad: float, iso=(0.0, 0.0, 0.0), ray_spacing=(1.0, 1.0), subspots=None) -> dict:
    energies = np.asarray(energies, dtype=np.float32)
    L = energies.size
    counts = np.array([len(w) for w in weights_per_layer], dtype=np.int32)
    pos = np.concatenate([np.asarray(p, dtype=np.float32).reshape(-1, 2) for p in spot_xy_per_layer])
    wts = np.concatenate([np.asarray(w, dtype=np.float32) for w in weights_per_layer])
    spot_sigmas = np.asarray(spot_sigmas, dtype=np.float32).reshape(L, 2)
    beam_dir = np.asarray(beam_dir, dtype=np.float64)
    src = np.asarray(iso, dtype=np.float64) - beam_dir * sad
    if subspots is None:
        subspots = single_subspot_lut(spot_sigmas)
    return {
        "energies": energies,
        "spot_sigmas": spot_sigmas,
        "ray_spacing": np.asarray(ray_spacing, dtype=np.float32),
        "steps": 0,
        "source_dist": np.array([sad, sad], dtype=np.float32),
        "spot_offset": np.zeros(3, np.float32),
        "spot_delta": np.zeros(3, np.float32),
        "beam_direction": beam_dir.astype(np.float32),
        "beam_xdir": np.asarray(bm_x, dtype=np.float32),
        "beam_ydir": np.asarray(bm_y, dtype=np.float32),
        "sad": np.float32(sad),
        "source_position": src.astype(np.float32),
        "ref_plane_z": np.float32(0.0),
        "layer_spot_counts": counts,
        "spot_positions": pos.ravel(),
        "spot_weights": wts,
        "subspot_data": np.ascontiguousarray(subspots, dtype=np.float32),
    }


def make_config1(seed: int = 1) -> dict:
    """Water phantom 64^3 at 2 mm with a 20 mm bone-like slab, one beam along +z.

    HU+1000 is supplied directly (water 1000, slab 2000): the reference's CT-type
    auto-detection reads a non-negative volume as HU+1000 (wrapper:2327-2335).
    5 layers 100..140 MeV, 11x11 spots at 5 mm, weights U[0.5,1.5], sigma 4 mm,
    a 5x5 subspot table (offsets -4..4 mm, sigma_sub 2 mm).
    """
    dims = (64, 64, 64)
    res = (2.0, 2.0, 2.0)
    corner = (-63.0, -63.0, -63.0)
    ct = np.full((dims[2], dims[1], dims[0]), 1000.0, dtype=np.float32)
    ct[20:30, :, :] = 2000.0
    energies = np.array([100.0, 110.0, 120.0, 130.0, 140.0], dtype=np.float32)
    g = np.arange(-25.0, 25.0 + 1e-6, 5.0)
    gx, gy = np.meshgrid(g, g, indexing="xy")
    xy = np.stack([gx.ravel(), gy.ravel()], axis=1)
    rng = Stream(seed, "config1.weights")
    weights = [0.5 + rng.uniform(xy.shape[0]) for _ in energies]
    sig = np.full((energies.size, 2), 4.0, dtype=np.float32)
    beam = make_beam(energies, [xy] * energies.size, weights, sig,
                     beam_dir=(0.0, 0.0, 1.0), bm_x=(1.0, 0.0, 0.0), bm_y=(0.0, 1.0, 0.0), sad=2000.0,
                     subspots=subspot_lut(energies.size, 2.0))
    return {
        "name": "config1-water-slab-64",
        "ct": ct, "ct_dims": dims, "ct_resolution": res, "ct_corner": corner,
        "dose_dims": dims, "dose_resolution": res, "dose_corner": corner,
        "beams": [beam], "energy": load_energy_tables(),
    }


def hex_lattice(pitch: float, radius: float):
    h = pitch * math.sqrt(3.0) / 2.0
    nmax = int(math.floor(radius / h)) + 1
    mmax = int(math.floor(radius / pitch)) + 2
    pts = []
    for n in range(-nmax, nmax + 1):
        y = n * h
        shift = 0.5 * pitch if (n % 2) else 0.0
        for m in range(-mmax, mmax + 1):
            x = m * pitch + shift
            if x * x + y * y <= radius * radius + 1e-9:
                pts.append((x, y))
    return np.array(pts, dtype=np.float64)


def head_phantom(seed: int = 2):
    """256x256x160 CT at 1x1x2 mm (HU): ellipsoidal body N(0,20), two bones U[700,1200],
    an air cavity, air outside."""
    dims = (256, 256, 160)
    res = (1.0, 1.0, 2.0)
    corner = (-127.5, -127.5, -159.0)
    xs, ys, zs = _grid_centres(dims, res, corner)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    body = (X / 80.0) ** 2 + (Y / 100.0) ** 2 + (Z / 70.0) ** 2 <= 1.0
    n = X.size
    ct = np.full(X.shape, -1000.0, dtype=np.float64)
    noise = Stream(seed, "head.body").normal(n).reshape(X.shape) * 20.0
    ct[body] = noise[body]
    ub = Stream(seed, "head.bone").uniform(n).reshape(X.shape)
    for cx in (-40.0, 40.0):
        bone = ((X - cx) / 12.0) ** 2 + ((Y - 20.0) / 15.0) ** 2 + (Z / 25.0) ** 2 <= 1.0
        ct[bone] = 700.0 + 500.0 * ub[bone]
    air = (X ** 2 + (Y + 35.0) ** 2 + (Z - 20.0) ** 2) <= 144.0
    ct[air] = -1000.0
    return ct.astype(np.float32), dims, res, corner


def head_beam(seed: int = 2, gantry_deg: float = 90.0, n_layers: int = 30,
              e_min: float = 70.0, e_max: float = 180.0, radius: float = 26.5) -> dict:
    """Head-size proton field: layers evenly spaced e_min..e_max MeV, 5 mm hexagonal spot
    lattice clipped to a disc, log-normal(0, 0.5) weights, sigma = 6 - 3 (E - 70) / 110 mm."""
    energies = np.linspace(e_min, e_max, n_layers).astype(np.float32)
    xy = hex_lattice(5.0, radius)
    rng = Stream(seed, f"head.weights.{gantry_deg:.3f}")
    weights = [np.exp(0.5 * rng.normal(xy.shape[0])) for _ in energies]
    s = (6.0 - 3.0 * (energies.astype(np.float64) - 70.0) / 110.0).astype(np.float32)
    sig = np.stack([s, s], axis=1)
    bdir, bmx, bmy = beam_basis(gantry_deg)
    return make_beam(energies, [xy] * energies.size, weights, sig,
                     beam_dir=bdir, bm_x=bmx, bm_y=bmy, sad=2000.0)


def make_config4(seed: int = 1) -> dict:
    """Config-1 beam with a weight-optimisation objective (differentiation case).

    Target: the 60 mm cube centred on the iso; OAR: an ellipsoid at the lateral edge of the
    field (centre (35, 0, 0) mm, semi-axes 12, 18, 25 mm). Objective
    J = sum_T (D - 2)^2 + 0.1 sum_OAR D^2, with the weights first scaled so that the mean target
    dose is 2 (the dose is linear in the weights). ``fd_spots``: 64 seeded spot indices for the
    central finite-difference check (h = 1e-3 w).
    """
    case = make_config1(seed)
    nx, ny, nz = case["dose_dims"]
    cx, cy, cz = case["dose_corner"]
    rx_, ry_, rz_ = case["dose_resolution"]
    z, y, x = np.meshgrid(cz + rz_ * np.arange(nz), cy + ry_ * np.arange(ny), cx + rx_ * np.arange(nx),
                          indexing="ij")
    case["target_mask"] = (np.abs(x) <= 30.0) & (np.abs(y) <= 30.0) & (np.abs(z) <= 30.0)
    case["oar_mask"] = ((x - 35.0) / 12.0) ** 2 + (y / 18.0) ** 2 + (z / 25.0) ** 2 <= 1.0
    n = int(np.asarray(case["beams"][0]["spot_weights"]).size)
    rng = Stream(seed, "config4.fd_spots")
    order = np.argsort(rng.uniform(n))
    case["fd_spots"] = np.sort(order[:64]).astype(np.int64)
    case["prescription"] = 2.0
    case["oar_factor"] = 0.1
    case["name"] = "config4-objective-64"
    return case


def scaled_head_phantom(dims, res, seed: int):
    """Head-like phantom on any grid: the config-2 body, bones and air cavity scaled to the
    grid's field of view (HU). Noise from numpy's PCG64 (large grids)."""
    corner = tuple(-0.5 * res[i] * (dims[i] - 1) for i in range(3))
    xs, ys, zs = _grid_centres(dims, res, corner)
    fov = [res[i] * dims[i] for i in range(3)]
    s = min(fov[0] / 256.0, fov[1] / 256.0, fov[2] / 320.0)
    rng = np.random.default_rng(seed)
    ct = np.full((dims[2], dims[1], dims[0]), -1000.0, dtype=np.float32)
    for k, z in enumerate(zs):
        Y, X = np.meshgrid(ys, xs, indexing="ij")
        body = (X / (80.0 * s)) ** 2 + (Y / (100.0 * s)) ** 2 + (z / (70.0 * s)) ** 2 <= 1.0
        sl = ct[k]
        sl[body] = (rng.standard_normal(int(body.sum())) * 20.0).astype(np.float32)
        for cx in (-40.0 * s, 40.0 * s):
            bone = ((X - cx) / (12.0 * s)) ** 2 + ((Y - 20.0 * s) / (15.0 * s)) ** 2 + (z / (25.0 * s)) ** 2 <= 1.0
            sl[bone] = (700.0 + 500.0 * rng.uniform(size=int(bone.sum()))).astype(np.float32)
        air = X ** 2 + (Y + 35.0 * s) ** 2 + (z - 20.0 * s) ** 2 <= (12.0 * s) ** 2
        sl[air] = -1000.0
    return ct, tuple(dims), tuple(res), corner


def make_config5(seed: int = 5, n_requests: int = 48, weight_only_fraction: float = 0.3) -> dict:
    """Serving mix: CTs 128^3 (2 mm), 256x256x160 (1x1x2 mm) and 512x512x200 (0.5x0.5x1.25 mm,
    dose on a 1x1x2.5 mm grid); each request one proton beam at a random gantry angle with
    500-8000 spots over 10-40 layers (5 mm hexagonal lattice), 70-180 MeV; a fraction of the
    requests are weight-only updates of an earlier request's beam (log-normal re-weighting).

    ``requests``: dicts with ``ct`` (index into ``cts``), ``beam``, and ``update_of`` (index of
    the request whose beam is re-weighted, or -1).
    """
    grids = [((128, 128, 128), (2.0, 2.0, 2.0), None), ((256, 256, 160), (1.0, 1.0, 2.0), None),
             ((512, 512, 200), (0.5, 0.5, 1.25), ((256, 256, 100), (1.0, 1.0, 2.5)))]
    cts = []
    for i, (dims, res, dose) in enumerate(grids):
        ct, dims, res, corner = scaled_head_phantom(dims, res, seed * 10 + i)
        if dose is None:
            dd, dr, dc = dims, res, corner
        else:
            dd, dr = dose
            dc = tuple(-0.5 * dr[j] * (dd[j] - 1) for j in range(3))
        cts.append({"ct": ct, "ct_dims": dims, "ct_resolution": res, "ct_corner": corner,
                    "dose_dims": tuple(dd), "dose_resolution": tuple(dr), "dose_corner": tuple(dc)})
    rng = Stream(seed, "config5.requests")
    u = rng.uniform(8 * n_requests)
    reqs = []
    for i in range(n_requests):
        a = u[8 * i: 8 * i + 8]
        full = [j for j, r in enumerate(reqs) if r["update_of"] < 0]
        if full and a[0] < weight_only_fraction:
            base = full[int(a[1] * len(full)) % len(full)]
            b = dict(reqs[base]["beam"])
            w = np.asarray(b["spot_weights"], np.float64)
            wr = Stream(seed, f"config5.reweight.{i}").normal(w.size)
            b["spot_weights"] = (w * np.exp(0.3 * wr)).astype(np.float32)
            reqs.append({"ct": reqs[base]["ct"], "beam": b, "update_of": base})
            continue
        n_spots = 500 + int(a[2] * 7501)
        n_layers = 10 + int(a[3] * 31)
        per_layer = max(1, n_spots // n_layers)
        radius = math.sqrt(per_layer * 25.0 * math.sqrt(3.0) / 2.0 / math.pi)
        xy = hex_lattice(5.0, radius)
        energies = np.linspace(70.0, 180.0, n_layers).astype(np.float32)
        wr = Stream(seed, f"config5.weights.{i}")
        weights = [np.exp(0.5 * wr.normal(xy.shape[0])) for _ in energies]
        sg = (6.0 - 3.0 * (energies.astype(np.float64) - 70.0) / 110.0).astype(np.float32)
        bdir, bmx, bmy = beam_basis(360.0 * float(a[4]))
        beam = make_beam(energies, [xy] * n_layers, weights, np.stack([sg, sg], axis=1), beam_dir=bdir, bm_x=bmx,
                         bm_y=bmy, sad=2000.0)
        reqs.append({"ct": int(a[5] * 3) % 3, "beam": beam, "update_of": -1})
    return {"name": "config5-serving-mix", "cts": cts, "requests": reqs, "energy": load_energy_tables()}


def _mixed_beam(rng_u, seed: int, tag: str, n_spots: int, n_layers: int, e_lo: float, e_hi: float,
                sigma_hi: float, sigma_lo: float, gantry_deg: float) -> dict:
    """A beam of ~n_spots spots over n_layers layers (hexagonal lattice, 5 mm pitch, tightened so the
    field stays within a 45 mm radius), log-normal(0, 0.5) weights, sigma falling linearly from
    sigma_hi at e_lo to sigma_lo at e_hi."""
    per_layer = max(1, n_spots // n_layers)
    radius = math.sqrt(per_layer * 25.0 * math.sqrt(3.0) / 2.0 / math.pi)
    pitch = 5.0 if radius <= 45.0 else 5.0 * 45.0 / radius
    xy = hex_lattice(pitch, min(radius, 45.0))
    energies = np.linspace(e_lo, e_hi, n_layers).astype(np.float32)
    wr = Stream(seed, tag)
    weights = [np.exp(0.5 * wr.normal(xy.shape[0])) for _ in energies]
    t = (energies.astype(np.float64) - e_lo) / max(e_hi - e_lo, 1e-9)
    sg = (sigma_hi + (sigma_lo - sigma_hi) * t).astype(np.float32)
    bdir, bmx, bmy = beam_basis(gantry_deg)
    return make_beam(energies, [xy] * n_layers, weights, np.stack([sg, sg], axis=1), beam_dir=bdir, bm_x=bmx,
                     bm_y=bmy, sad=2000.0)


def make_config3(seed: int = 3, n_proton: int = 12, n_carbon: int = 4) -> dict:
    """Batching case on the config-1 CT: proton beams at gantry angles uniform in [0, 360) with
    500-8000 spots over 10-40 layers (70-150 MeV, sigma 6 -> 3 mm) and carbon-ion beams of
    20 layers within 100-400 MeV/u (500-2000 spots, sigma 4 -> 2.5 mm) on the project-defined
    carbon tables (``subsecondpbs_amd.carbon``, not reference data) with the halo branch.

    Keys: ``beams`` (protons then carbons), ``kinds`` ("proton" / "carbon" per beam), ``energy``
    (proton tables), ``carbon_energy`` (carbon tables with halo matrices), ``carbon_nuclear``
    (halo model of the carbon beams: 3, FLUKA constants).
    """
    from .carbon import build_carbon_tables

    case = make_config1(seed)
    rng = Stream(seed, "config3.beams")
    beams, kinds = [], []
    for i in range(n_proton):
        a = rng.uniform(4)
        beams.append(_mixed_beam(a, seed, f"config3.proton.{i}", 500 + int(a[0] * 7501), 10 + int(a[1] * 31),
                                 70.0, 150.0, 6.0, 3.0, 360.0 * float(a[2])))
        kinds.append("proton")
    for i in range(n_carbon):
        a = rng.uniform(4)
        e_lo = 100.0 + 100.0 * float(a[1])
        beams.append(_mixed_beam(a, seed, f"config3.carbon.{i}", 500 + int(a[0] * 1501), 20,
                                 e_lo, e_lo + 200.0 * (0.5 + 0.5 * float(a[3])), 4.0, 2.5, 360.0 * float(a[2])))
        kinds.append("carbon")
    case.update({"name": "config3-batch-16-mixed", "beams": beams, "kinds": kinds,
                 "carbon_energy": build_carbon_tables(), "carbon_nuclear": 3})
    return case


def case_for_beam(case: dict, i: int) -> dict:
    """A one-beam case (CT, grids, tables) for beam ``i`` of a mixed-ion case."""
    out = dict(case)
    out["beams"] = [case["beams"][i]]
    if case.get("kinds") and case["kinds"][i] == "carbon":
        out["energy"] = case["carbon_energy"]
    return out


def weq_volume(header, density_fn) -> np.ndarray:
    """WEQ vector (header + samples [x][y][k], raytracedicom_integration.h:63-70) with
    WEPL(x, y, k) = step x sum_{j <= k} density_fn(ix, iy, j)."""
    h = np.asarray(header, dtype=np.float64)
    nk, ny, nx = int(round(h[2])), int(round(h[5])), int(round(h[8]))
    ix, iy, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nk), indexing="ij")
    d = density_fn(ix, iy, k)
    body = np.cumsum(d, axis=2) * h[1]
    return np.concatenate([h, body.ravel()]).astype(np.float32)


def make_weq_case(seed: int = 6, index_spots: bool = False, start_shift_mm: float = 4.0) -> dict:
    """WEQ-volume input (the CarbonPBS contract): the config-1 CT bounds the geometry, but the
    beam carries its own water-equivalent depths on an 81 x 71 lattice at 1 mm, 64 steps of 2 mm:
    water-like density 1 + 0.01 ix (a lateral wedge along x only, so a transposed [x][y] index
    shows), plus 0.5 in steps 20-29 for y >= 35 (a half slab). depth_start is set so the start
    plane moves by ``start_shift_mm`` from the CT's. Spots: the config-1 11 x 11 field, 5 layers,
    at physical positions or (``index_spots``) as WEQ texel coordinates (+0.5 at texel centres),
    the corner spot of every layer a quarter texel off so that it splits over four cells."""
    case = make_config1(seed)
    beam = dict(case["beams"][0])
    # gantry z of a world point is -z (origin at the isocentre), so the CT's start plane is at
    # gantry z = 63; depth_start = (mean source distance) - (the start plane wanted)
    depth_start = 2000.0 - (63.0 + start_shift_mm)
    header = [depth_start, 2.0, 64.0, -35.0, 1.0, 71.0, -40.0, 1.0, 81.0]
    slab = lambda ix, iy, k: 0.5 * ((k >= 20) & (k < 30) & (iy >= 35))
    beam["water_equivalence"] = weq_volume(header, lambda ix, iy, k: 1.0 + 0.01 * ix + slab(ix, iy, k))
    if index_spots:
        pos = np.asarray(beam["spot_positions"], np.float32).reshape(-1, 2).copy()
        raw = np.empty_like(pos)
        raw[:, 0] = (pos[:, 0] - header[6]) / header[7] + 0.5
        raw[:, 1] = (pos[:, 1] - header[3]) / header[4] + 0.5
        per = len(pos) // len(beam["energies"])
        raw[::per] += 0.25  # the first spot of each layer: between texel centres
        beam["spot_positions"] = raw.ravel()
        beam["spot_positions_are_indices"] = True
    case["beams"] = [beam]
    case["name"] = "weq-volume-64" + ("-index-spots" if index_spots else "")
    case["weq_header"] = header
    return case


def make_config2(seed: int = 2, gantry_deg: float = 90.0) -> dict:
    """Head-size phantom with one 30-layer proton beam at gantry 90 deg (along +x)."""
    ct, dims, res, corner = head_phantom(seed)
    return {
        "name": "config2-head-256x256x160-30L",
        "ct": ct, "ct_dims": dims, "ct_resolution": res, "ct_corner": corner,
        "dose_dims": dims, "dose_resolution": res, "dose_corner": corner,
        "beams": [head_beam(seed, gantry_deg)], "energy": load_energy_tables(),
    }


def make_plan_4field(seed: int = 2) -> dict:
    """Config-2 phantom with four coplanar fields (gantry 0/90/180/270): the <100 ms plan."""
    base = make_config2(seed)
    base["beams"] = [head_beam(seed, g) for g in (0.0, 90.0, 180.0, 270.0)]
    base["name"] = "plan-4field-head"
    return base
