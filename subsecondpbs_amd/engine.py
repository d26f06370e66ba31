"""Resident dose engine: CT, LUTs and work buffers stay in HBM across beams.

Thin Python handle over the ``pbs_engine_*`` C ABI. Device buffers passed in (dose grids,
CT) are torch tensors on the HIP device; the engine queues work on its own stream, so
callers synchronise with :meth:`Engine.sync` before reading results.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _abi


def torch_ready():
    """Block until the work queued so far on torch's current stream has run. For device tensors
    handed to native streams other than an engine's (batch lanes, server copies)."""
    import torch

    torch.cuda.current_stream().synchronize()


class PreparedBeam:
    """A beam's ``pbs_beam_settings`` struct, built once from the beam dict.

    This is the struct a C or C++ caller of ``pbs_engine_beam_dose`` holds itself. Every call with
    it still plans, uploads and computes the beam; only the dict -> struct conversion is not
    repeated. The struct views the beam's arrays where they are contiguous and of the struct's
    type (and copies otherwise), so prepare again after changing the beam."""

    def __init__(self, beam: dict, use_subspots: bool | None = None):
        self._keep = _abi.Keep()
        self.settings = _abi.beam_settings(beam, self._keep, use_subspots)
        self.num_spots = int(self.settings.numSpots)


class Engine:
    def __init__(self, energy: dict, device: int = 0):
        self._keep = _abi.Keep()
        self._es = _abi.energy_struct(energy, self._keep)
        h = _abi.lib().pbs_engine_create(int(device), C.byref(self._es))
        if not h:
            raise RuntimeError("pbs_engine_create failed: " + _abi.last_error())
        self.handle = h
        self.device = device
        self.ct_dims = None
        self.dose_dims = None
        self._resident = weakref.WeakSet()
        self._torch_ev = None
        self._torch_stream = None

    def after_torch(self):
        """Order this engine's stream after the work queued so far on torch's current stream, on
        the device (an event, no host wait): a dose tensor that torch has just zeroed or written
        is then safe to accumulate into. Every call below that takes a torch tensor does this."""
        import torch

        cur = torch.cuda.current_stream()
        if cur.query():
            return  # everything torch queued has run: no device-side wait needed
        if self._torch_ev is None:
            self._torch_ev = torch.cuda.Event()
            self._torch_stream = torch.cuda.ExternalStream(self.stream_ptr)
        self._torch_ev.record(cur)
        self._torch_stream.wait_event(self._torch_ev)

    def close(self):
        if getattr(self, "handle", None):
            for rb in list(getattr(self, "_resident", ())):
                rb.close()  # resident beams hold device state of this engine
            _abi.lib().pbs_engine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ geometry
    def set_ct(self, ct, dims, resolution, corner):
        """ct: numpy float32 [nz][ny][nx] (host) or a torch tensor on the device."""
        d, dp = _abi.i3(dims)
        r, rp = _abi.f3(resolution)
        c, cp = _abi.f3(corner)
        if isinstance(ct, np.ndarray):
            a = np.ascontiguousarray(ct, dtype=np.float32)
            rc = _abi.lib().pbs_engine_set_ct(self.handle, a.ctypes.data, dp, rp, cp, 0)
        else:
            torch_ready()  # the engine copies the CT synchronously
            rc = _abi.lib().pbs_engine_set_ct(self.handle, C.c_void_p(ct.data_ptr()), dp, rp, cp, 1)
        _abi.check(rc, "pbs_engine_set_ct")
        self.ct_dims = tuple(int(v) for v in dims)

    def set_dose_grid(self, dims, resolution, corner):
        d, dp = _abi.i3(dims)
        r, rp = _abi.f3(resolution)
        c, cp = _abi.f3(corner)
        _abi.check(_abi.lib().pbs_engine_set_dose_grid(self.handle, dp, rp, cp), "pbs_engine_set_dose_grid")
        self.dose_dims = tuple(int(v) for v in dims)

    # ------------------------------------------------------------------ compute
    @staticmethod
    def prepare(beam: dict, use_subspots: bool | None = None) -> PreparedBeam:
        """The beam's C struct, for repeated :meth:`beam_dose` calls (see PreparedBeam)."""
        return PreparedBeam(beam, use_subspots)

    def beam_dose(self, beam, dose, use_subspots: bool | None = None):
        """Queue one beam (a dict or a PreparedBeam); its dose is added into ``dose`` (device
        tensor [nz][ny][nx] float32)."""
        if isinstance(beam, PreparedBeam):
            bs = beam.settings
        else:
            keep = _abi.Keep()
            bs = _abi.beam_settings(beam, keep, use_subspots)
        self.after_torch()
        rc = _abi.lib().pbs_engine_beam_dose(self.handle, C.byref(bs), C.c_void_p(dose.data_ptr()))
        _abi.check(rc, "pbs_engine_beam_dose")

    def beam_adjoint(self, beam: dict, grad_dose, use_subspots: bool | None = None) -> np.ndarray:
        """dJ/d(spot weights) of one beam for dJ/dDose = ``grad_dose`` (device tensor
        [nz][ny][nx] float32), at the beam's weights; float32 numpy [numSpots]. Synchronous."""
        keep = _abi.Keep()
        bs = _abi.beam_settings(beam, keep, use_subspots)
        out = np.zeros(int(bs.numSpots), np.float32)
        self.after_torch()
        rc = _abi.lib().pbs_engine_beam_adjoint(self.handle, C.byref(bs), C.c_void_p(grad_dose.data_ptr()),
                                                C.c_void_p(out.ctypes.data))
        _abi.check(rc, "pbs_engine_beam_adjoint")
        return out

    def keep_beam(self, beam: dict, dose, use_subspots: bool | None = None) -> "ResidentBeam":
        """Queue the beam into ``dose`` and keep its weight-independent state resident on this
        engine; weight-only updates and gradients then skip the trace and the fill."""
        keep = _abi.Keep()
        bs = _abi.beam_settings(beam, keep, use_subspots)
        h = C.c_void_p(0)
        self.after_torch()
        rc = _abi.lib().pbs_engine_keep_beam(self.handle, C.byref(bs), C.c_void_p(dose.data_ptr()), C.byref(h))
        _abi.check(rc, "pbs_engine_keep_beam")
        rb = ResidentBeam(self, h.value, int(bs.numSpots))
        self._resident.add(rb)
        return rb

    def sync(self):
        _abi.check(_abi.lib().pbs_engine_sync(self.handle), "pbs_engine_sync")

    STAGES = ("trace", "conv", "fill", "tiles", "superpose", "merge", "transfer", "halo")

    NUCLEAR_MODELS = {"off": 0, "gauss_fit": 1, "soukup": 2, "fluka": 3}

    def set_nuclear(self, model: int | str | bool):
        """Halo branch for the following beams: 0/False off, 1/True GAUSS_FIT, 2 SOUKUP, 3 FLUKA."""
        if isinstance(model, str):
            model = self.NUCLEAR_MODELS[model.lower()]
        _abi.check(_abi.lib().pbs_engine_set_nuclear(self.handle, int(model)), "pbs_engine_set_nuclear")

    def debug_halo(self) -> dict:
        """The last beam's halo grid: dims, spotDist, (IDD, rSigmaEff) fields, radius histograms."""
        dims = (C.c_int * 2)()
        sd = C.c_float(0.0)
        rc = _abi.lib().pbs_engine_debug_halo(self.handle, dims, C.byref(sd), None, None, None)
        _abi.check(rc, "pbs_engine_debug_halo")
        ints = self.debug_fetch(())["ints"]
        L, K = ints["numLayers"], ints["steps"]
        nx, ny = dims[0], dims[1]
        f = np.zeros((L, K, ny, nx, 2), np.float32)
        h = np.zeros((L, 34), np.int32)
        ov = C.c_int(0)
        rc = _abi.lib().pbs_engine_debug_halo(self.handle, dims, C.byref(sd), C.c_void_p(f.ctypes.data),
                                              C.c_void_p(h.ctypes.data), C.byref(ov))
        _abi.check(rc, "pbs_engine_debug_halo")
        return {"nucDims": (nx, ny), "spotDist": float(sd.value), "iddSigma": f, "radHist": h,
                "overflow": int(ov.value)}

    TUNING_KEYS = {"superpose_grid": 1, "transfer_shape": 2, "one_stream": 3}

    def set_tuning(self, **kw):
        """Launch shapes (see ``subsecondpbs_amd.autotune``): superpose_grid, transfer_shape,
        one_stream. They change timing only, never a result."""
        for k, v in kw.items():
            _abi.check(_abi.lib().pbs_engine_set_tuning(self.handle, self.TUNING_KEYS[k], int(v)),
                       "pbs_engine_set_tuning")

    def tuning(self) -> dict:
        out = {}
        for k, key in self.TUNING_KEYS.items():
            v = C.c_int(0)
            _abi.check(_abi.lib().pbs_engine_get_tuning(self.handle, key, C.byref(v)), "pbs_engine_get_tuning")
            out[k] = int(v.value)
        return out

    def debug_bins(self, halo: bool = False) -> np.ndarray:
        """The last beam's batched launch radius per radius bin, [numLayers][34] (halo grid if
        ``halo``)."""
        L = self.debug_fetch(())["ints"]["numLayers"]
        out = np.zeros((L, 34), np.int32)
        ptr = C.c_void_p(out.ctypes.data)
        rc = _abi.lib().pbs_engine_debug_bins(self.handle, None if halo else ptr, ptr if halo else None)
        _abi.check(rc, "pbs_engine_debug_bins")
        return out

    def set_profiling(self, on: bool = True):
        _abi.check(_abi.lib().pbs_engine_set_profiling(self.handle, int(bool(on))), "pbs_engine_set_profiling")

    def stage_times(self) -> tuple[dict, int]:
        """Device ms per stage summed over the beams queued since the last call, and their count."""
        ms = (C.c_float * len(self.STAGES))()
        n = C.c_int(0)
        _abi.check(_abi.lib().pbs_engine_stage_times(self.handle, ms, C.byref(n)), "pbs_engine_stage_times")
        return {k: float(v) for k, v in zip(self.STAGES, ms)}, int(n.value)

    @property
    def stream_ptr(self) -> int:
        return int(_abi.lib().pbs_engine_stream(self.handle) or 0)

    def debug_fetch(self, what=("density", "wepl", "fi", "fo", "weplMin", "rayW", "iddSigma", "firstPassive",
                                 "tileRad", "radHist", "bev")) -> dict:
        """Copy the last beam's intermediate arrays to host numpy arrays."""
        ints = _abi.DebugInts()
        rc = _abi.lib().pbs_engine_debug_fetch(self.handle, C.byref(ints), *([None] * 11))
        _abi.check(rc, "pbs_engine_debug_fetch")
        rx, ry = ints.rayDims
        sx, sy = ints.superpDims
        K, L = ints.steps, ints.numLayers
        nt = (sx // 32) * (sy // 8)
        shapes = {
            "density": ((K, ry, rx), np.float32), "wepl": ((K, ry, rx), np.float32),
            "fi": ((ry, rx), np.int32), "fo": ((ry, rx), np.int32), "weplMin": ((K,), np.float32),
            "rayW": ((L, ry, rx), np.float32), "iddSigma": ((L, K, sy, sx, 2), np.float32),
            "firstPassive": ((L, ry, rx), np.int32), "tileRad": ((L, K, sy // 8, sx // 32), np.int32),
            "radHist": ((L, 34), np.int32), "bev": ((K, sy + 64, sx + 64), np.float32),
        }
        order = ["density", "wepl", "fi", "fo", "weplMin", "rayW", "iddSigma", "firstPassive", "tileRad",
                 "radHist", "bev"]
        out = {}
        ptrs = []
        for name in order:
            if name in what:
                shp, dt = shapes[name]
                a = np.zeros(shp, dt)
                out[name] = a
                ptrs.append(C.c_void_p(a.ctypes.data))
            else:
                ptrs.append(None)
        rc = _abi.lib().pbs_engine_debug_fetch(self.handle, C.byref(ints), *ptrs)
        _abi.check(rc, "pbs_engine_debug_fetch")
        out["ints"] = {
            "rayDims": (rx, ry), "superpDims": (sx, sy), "steps": K, "numLayers": L, "legacy": ints.legacyPath,
            "beamFirstInside": ints.beamFirstInside, "beamFirstOutside": ints.beamFirstOutside,
            "firstPastCutoffAll": ints.firstPastCutoffAll, "guaranteedPassive": ints.guaranteedPassive,
            "afterLast": np.array(ints.afterLast[:L], np.int32), "layerPassive": np.array(ints.layerPassive[:L], np.int32),
            "transferStart": tuple(ints.transferStart), "transferMax": tuple(ints.transferMax),
            "cpbCorner": tuple(ints.cpbCorner), "fanCorner": tuple(ints.fanCorner), "fanDelta": tuple(ints.fanDelta),
            "sourceDist": tuple(ints.sourceDist),
            "energyIdx": np.array(ints.energyIdx[:L], np.float32), "scaleFact": np.array(ints.scaleFact[:L], np.float32),
            "peakDepth": np.array(ints.peakDepth[:L], np.float32),
            "numGroups": ints.numGroups, "tailPlanes": ints.tailPlanes, "tailOverflow": ints.tailOverflow,
            "mainEndMax": ints.mainEndMax,
            "mainEnd": np.array(ints.mainEnd[:L], np.int32), "layerGroup": np.array(ints.layerGroup[:L], np.int32),
            "groupPassive": np.array(ints.groupPassive[:ints.numGroups], np.int32),
            "groupTailStart": np.array(ints.groupTailStart[:ints.numGroups], np.int32),
            "groupTailCount": np.array(ints.groupTailCount[:ints.numGroups], np.int32),
        }
        return out


class ResidentBeam:
    """A beam kept resident on an Engine (``Engine.keep_beam``): its unit-weight IDD / sigma
    field, plan and geometry stay in HBM. ``update`` adds the dose at new spot weights to a device
    buffer; ``adjoint`` gives dJ/dw at new weights; ``use_graph(True)`` replays updates as one
    captured HIP graph (while the engine, dose buffer and device buffers stay the same)."""

    def __init__(self, engine, handle, num_spots: int):
        self.engine = engine
        self.handle = handle
        self.num_spots = num_spots

    def update(self, weights, dose):
        w = np.ascontiguousarray(weights, dtype=np.float32)
        self.engine.after_torch()
        rc = _abi.lib().pbs_engine_update_beam(self.engine.handle, self.handle, w.ctypes.data_as(C.POINTER(C.c_float)),
                                               int(w.size), C.c_void_p(dose.data_ptr()))
        _abi.check(rc, "pbs_engine_update_beam")

    def dose(self, weights):
        """The beam's dose at ``weights`` as a new device tensor (Z, Y, X). Computed into a buffer
        kept by this object, so a graph captured for that buffer is replayed call after call."""
        import torch

        nx, ny, nz = self.engine.dose_dims
        buf = getattr(self, "_buf", None)
        if buf is None or tuple(buf.shape) != (nz, ny, nx):
            buf = self._buf = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
        else:
            buf.zero_()
        torch.cuda.current_stream().synchronize()  # the engine stream reads it next
        self.update(weights, buf)
        self.engine.sync()
        return buf.clone()

    def adjoint(self, weights, grad_dose) -> np.ndarray:
        w = np.ascontiguousarray(weights, dtype=np.float32)
        out = np.zeros(self.num_spots, np.float32)
        self.engine.after_torch()
        rc = _abi.lib().pbs_engine_update_adjoint(self.engine.handle, self.handle,
                                                  w.ctypes.data_as(C.POINTER(C.c_float)), int(w.size),
                                                  C.c_void_p(grad_dose.data_ptr()), C.c_void_p(out.ctypes.data))
        _abi.check(rc, "pbs_engine_update_adjoint")
        return out

    def use_graph(self, on: bool = True):
        _abi.check(_abi.lib().pbs_resident_beam_use_graph(self.handle, int(bool(on))), "pbs_resident_beam_use_graph")

    def close(self):
        if self.handle:
            _abi.lib().pbs_resident_beam_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
