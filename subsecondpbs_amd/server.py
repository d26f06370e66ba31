"""Served form of the final-dose engine: a request server over engine slots.

``Server`` wraps the ``pbs_server_*`` C ABI. CTs are resident (``add_ct``); ``submit`` queues a
beam dose request and returns a ticket at once; a native worker thread admits queued requests
onto free engine slots while the device memory they need fits the budget, so requests overlap on
the GPU (continuous batching). ``submit(..., keep=True)`` keeps the beam resident; then
``submit_weights(handle, weights)`` re-runs only the convolution, the rescale of the resident
unit-weight field, the superposition, the merge and the transfer. ``result`` waits and returns
the dose as float32 numpy (Z, Y, X); ``timing`` gives the queue and total latency in ms.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .engine import torch_ready


class Server:
    def __init__(self, energy: dict, slots: int = 4, device: int = 0, memory_budget_bytes: float = 0.0):
        self._keep = _abi.Keep()
        self._es = _abi.energy_struct(energy, self._keep)
        h = _abi.lib().pbs_server_create(int(device), C.byref(self._es), int(slots), float(memory_budget_bytes))
        if not h:
            raise RuntimeError("pbs_server_create: " + _abi.last_error())
        self.handle = h
        self.slots = int(slots)
        self._dims = {}  # ticket -> dose dims (x, y, z)
        self._state_dims = {}

    def close(self):
        if getattr(self, "handle", None):
            _abi.lib().pbs_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_ct(self, ct, dims, resolution, corner) -> int:
        d, dp = _abi.i3(dims)
        r, rp = _abi.f3(resolution)
        c, cp = _abi.f3(corner)
        cid = C.c_int(0)
        if isinstance(ct, np.ndarray):
            a = np.ascontiguousarray(ct, dtype=np.float32)
            rc = _abi.lib().pbs_server_add_ct(self.handle, C.c_void_p(a.ctypes.data), dp, rp, cp, 0, C.byref(cid))
        else:
            torch_ready()  # the server copies the CT from it
            rc = _abi.lib().pbs_server_add_ct(self.handle, C.c_void_p(ct.data_ptr()), dp, rp, cp, 1, C.byref(cid))
        _abi.check(rc, "pbs_server_add_ct")
        return int(cid.value)

    def remove_ct(self, ct_id: int):
        _abi.check(_abi.lib().pbs_server_remove_ct(self.handle, int(ct_id)), "pbs_server_remove_ct")

    def submit(self, ct_id: int, dose_dims, dose_resolution, dose_corner, beam: dict, keep: bool = False,
               use_subspots: bool | None = None) -> int:
        k = _abi.Keep()
        bs = _abi.beam_settings(beam, k, use_subspots)
        d, dp = _abi.i3(dose_dims)
        r, rp = _abi.f3(dose_resolution)
        c, cp = _abi.f3(dose_corner)
        t = C.c_longlong(0)
        rc = _abi.lib().pbs_server_submit(self.handle, int(ct_id), dp, rp, cp, C.byref(bs), int(bool(keep)),
                                          C.byref(t))
        _abi.check(rc, "pbs_server_submit")
        self._dims[int(t.value)] = tuple(int(v) for v in dose_dims)
        if keep:
            self._state_dims[int(t.value)] = self._dims[int(t.value)]
        return int(t.value)

    def submit_weights(self, handle: int, weights) -> int:
        w = np.ascontiguousarray(weights, dtype=np.float32)
        t = C.c_longlong(0)
        rc = _abi.lib().pbs_server_submit_weights(self.handle, int(handle), w.ctypes.data_as(C.POINTER(C.c_float)),
                                                  int(w.size), C.byref(t))
        _abi.check(rc, "pbs_server_submit_weights")
        self._dims[int(t.value)] = self._state_dims[int(handle)]
        return int(t.value)

    def query(self, ticket: int) -> bool:
        rc = _abi.lib().pbs_server_query(self.handle, int(ticket))
        if rc < 0:
            _abi.check(rc, "pbs_server_query")
        return rc == 1

    def wait(self, ticket: int):
        _abi.check(_abi.lib().pbs_server_wait(self.handle, int(ticket)), "pbs_server_wait")

    def result(self, ticket: int, device: bool = False):
        """The dose as float32 (Z, Y, X): numpy on the host, or (device=True) a torch tensor on
        the GPU (no host transfer)."""
        nx, ny, nz = self._dims.pop(int(ticket))
        if device:
            import torch

            out = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
            torch_ready()  # the server's copy stream writes it next
            rc = _abi.lib().pbs_server_result_device(self.handle, int(ticket), C.c_void_p(out.data_ptr()))
            _abi.check(rc, "pbs_server_result_device")
            return out
        out = np.empty((nz, ny, nx), np.float32)
        rc = _abi.lib().pbs_server_result(self.handle, int(ticket), C.c_void_p(out.ctypes.data))
        _abi.check(rc, "pbs_server_result")
        return out

    def timing(self, ticket: int) -> tuple[float, float]:
        q, t = C.c_double(0), C.c_double(0)
        _abi.check(_abi.lib().pbs_server_timing(self.handle, int(ticket), C.byref(q), C.byref(t)), "pbs_server_timing")
        return float(q.value), float(t.value)

    def release(self, handle: int):
        _abi.check(_abi.lib().pbs_server_release_state(self.handle, int(handle)), "pbs_server_release_state")
        self._state_dims.pop(int(handle), None)

    def memory(self) -> tuple[float, float]:
        u, b = C.c_double(0), C.c_double(0)
        _abi.check(_abi.lib().pbs_server_memory(self.handle, C.byref(u), C.byref(b)), "pbs_server_memory")
        return float(u.value), float(b.value)
