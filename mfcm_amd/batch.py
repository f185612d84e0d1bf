"""Batches of catchments advanced together (the mfsv_batch_* C ABI).

A ``Batch`` references ``Catchment`` members that keep their own device
fields and controllers; ``run`` advances every unfinished member by a number
of attempted timesteps, one stage launch per RK stage for all of
them.  A member's result is the same alone (``Catchment.run``) or in a batch:
every tile runs the same kernel code on that member's data only.
"""
from __future__ import annotations

import ctypes as C

from .mfsv import ERRORS, MfsvError, lib


class Batch:
    def __init__(self, precision, device=0):
        h = C.c_void_p()
        rc = lib().mfsv_batch_create(precision, device, C.byref(h))
        if rc != 0:
            raise MfsvError(f"mfsv_batch_create failed ({rc})")
        self._h = h
        self.precision = precision
        self.members = []

    def _check(self, rc):
        if rc != 0:
            msg = lib().mfsv_batch_last_error(self._h).decode() if self._h else ""
            raise MfsvError(f"mfsv batch error {rc} ({ERRORS.get(rc, '?')}): {msg}")

    def add(self, cat):
        self._check(lib().mfsv_batch_add(self._h, cat._h))
        self.members.append(cat)

    def remove(self, cat):
        self._check(lib().mfsv_batch_remove(self._h, cat._h))
        self.members.remove(cat)

    def size(self):
        m, a = C.c_int(), C.c_int()
        self._check(lib().mfsv_batch_size(self._h, C.byref(m), C.byref(a)))
        return m.value, a.value

    def run(self, max_timesteps=1 << 40, graph=True):
        self._check(lib().mfsv_batch_run(self._h, int(max_timesteps), 1 if graph else 0))
        return self.size()[1]

    def last_run_ms(self):
        m = C.c_double()
        self._check(lib().mfsv_batch_last_run_ms(self._h, C.byref(m)))
        return m.value

    def close(self):
        if getattr(self, "_h", None):
            lib().mfsv_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
