"""Batched and served forms of the final-dose engine on one GPU.

``Batch.dose`` runs a batch of beams as one multi-beam launch per pass of the chain (trace,
upload, setup, convolution, fill, tiles, each radius class, merge, and one transfer for all
beams that adds them voxel by voxel in beam order): the dose equals the one-beam-at-a-time sum
bitwise, and the launches do not grow with the beam count. Beams of several energy-table sets
(``add_table``: carbon beside protons) share the passes. A batch holding beams the multi-beam
passes do not cover runs on the lanes instead (own HIP stream and workspaces each, sharing the
resident CT; lane doses added in a fixed order). ``submit`` is the served form: a request with its own dose buffer gets the same result alone
and among others; requests into one buffer are queued on one lane and add up in order.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .engine import torch_ready


class Batch:
    def __init__(self, energy: dict, lanes: int = 4, device: int = 0):
        self._keep = _abi.Keep()
        es = _abi.energy_struct(energy, self._keep)
        self._es = es
        h = _abi.lib().pbs_batch_create(device, C.byref(es), int(lanes))
        if not h:
            raise RuntimeError("pbs_batch_create: " + _abi.last_error())
        self.handle = h
        self.lanes = int(lanes)
        self._beam_keep = {}

    def close(self):
        if getattr(self, "handle", None):
            _abi.lib().pbs_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_ct(self, ct, dims, resolution, corner):
        import torch
        if isinstance(ct, torch.Tensor):
            if not (ct.is_cuda and ct.dtype == torch.float32 and ct.is_contiguous()):
                raise ValueError("device CT must be a contiguous float32 tensor")
            torch_ready()  # the lanes copy the CT from it
            ptr, mode = C.c_void_p(ct.data_ptr()), 1
        else:
            a = np.ascontiguousarray(ct, dtype=np.float32)
            self._ct_host = a
            ptr, mode = C.c_void_p(a.ctypes.data), 0
        d, dp = _abi.i3(dims)
        r, rp = _abi.f3(resolution)
        c, cp = _abi.f3(corner)
        _abi.check(_abi.lib().pbs_batch_set_ct(self.handle, ptr, dp, rp, cp, mode), "pbs_batch_set_ct")

    def set_dose_grid(self, dims, resolution, corner):
        d, dp = _abi.i3(dims)
        r, rp = _abi.f3(resolution)
        c, cp = _abi.f3(corner)
        _abi.check(_abi.lib().pbs_batch_set_dose_grid(self.handle, dp, rp, cp), "pbs_batch_set_dose_grid")

    def set_nuclear(self, model: int):
        _abi.check(_abi.lib().pbs_batch_set_nuclear(self.handle, int(model)), "pbs_batch_set_nuclear")

    def add_table(self, energy: dict, nuclear: int = 0) -> int:
        """Another energy-table set (e.g. carbon beside protons) and its nuclear model; returns
        its id for ``dose(..., tables=...)`` (the set given at creation is 0)."""
        es = _abi.energy_struct(energy, self._keep)
        rc = _abi.lib().pbs_batch_add_table(self.handle, C.byref(es), int(nuclear))
        if rc < 0:
            _abi.check(rc, "pbs_batch_add_table")
        return int(rc)

    def prepare(self, beams: list[dict]):
        """The beams' C structs, built once (a caller that repeats a batch keeps them)."""
        keep = _abi.Keep()
        arr = (_abi.BeamSettings * len(beams))()
        for i, b in enumerate(beams):
            arr[i] = _abi.beam_settings(b, keep)
        return arr, keep

    def dose(self, beams, dose, tables=None):
        """Queue the sum of the beams' doses into ``dose`` (device tensor [nz][ny][nx]).
        ``beams``: beam dicts or the result of ``prepare``; ``tables``: each beam's table-set id."""
        arr, keep = beams if isinstance(beams, tuple) else self.prepare(beams)
        n = len(arr)
        torch_ready()  # the batch's streams read and write it next
        if tables is None:
            rc = _abi.lib().pbs_batch_dose(self.handle, arr, n, C.c_void_p(dose.data_ptr()))
            _abi.check(rc, "pbs_batch_dose")
        else:
            t = (C.c_int * n)(*[int(x) for x in tables])
            rc = _abi.lib().pbs_batch_dose_tables(self.handle, arr, t, n, C.c_void_p(dose.data_ptr()))
            _abi.check(rc, "pbs_batch_dose_tables")

    @property
    def graph_captures(self) -> int:
        """HIP graphs captured so far (a repeated batch shape replays its graph)."""
        return int(_abi.lib().pbs_batch_graph_captures(self.handle))

    @property
    def last_launches(self) -> int:
        """Kernel launches of the last multi-beam call (0 when the lanes ran it)."""
        return int(_abi.lib().pbs_batch_last_launches(self.handle))

    def submit(self, beam: dict, dose) -> int:
        """Queue one beam whose dose is added into ``dose``; returns a ticket. A buffer that a
        queued request still writes goes to that request's lane, so requests into one buffer
        accumulate in submission order."""
        keep = _abi.Keep()
        bs = _abi.beam_settings(beam, keep)
        t = C.c_longlong(0)
        torch_ready()  # the lanes' streams read and write it next
        _abi.check(_abi.lib().pbs_batch_submit(self.handle, C.byref(bs), C.c_void_p(dose.data_ptr()), C.byref(t)),
                   "pbs_batch_submit")
        return int(t.value)

    def query(self, ticket: int) -> bool:
        rc = _abi.lib().pbs_batch_query(self.handle, C.c_longlong(ticket))
        if rc < 0:
            _abi.check(rc, "pbs_batch_query")
        return rc == 1

    def wait(self, ticket: int):
        _abi.check(_abi.lib().pbs_batch_wait(self.handle, C.c_longlong(ticket)), "pbs_batch_wait")

    def sync(self):
        _abi.check(_abi.lib().pbs_batch_sync(self.handle), "pbs_batch_sync")
