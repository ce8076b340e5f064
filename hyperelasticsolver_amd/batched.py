"""Batched solves: thousands of independent 1D problems in one launch per step.

Each problem has its own length, materials, CFL number and dx, hence its own
time step (main.jl:204-214 applied per problem).  The state of all problems is
packed (ncomp x total cells, problems one after another) and lives on the GPU.
`run(steps, mode="graph")` captures the whole step loop in one hipGraph;
`mode="loop"` is the one-problem-per-launch baseline; `autotune()` picks the
launch shape by timing each candidate on the actual problem set.
"""
import ctypes

import numpy as np

from ._lib import check, columns, lib, ptr
from .solver import FLUXES, materials_array, ncomp_of, nphase_of

MODES = {"eager": 0, "graph": 1, "loop": 2}


class Batch:
    def __init__(self, eos_list, ncells, flux="rusanov", cfl=0.6, dx=None):
        """eos_list: one entry per problem (a Barton2009, or a tuple of two for
        two-phase problems); ncells: cells per problem."""
        self.P = len(eos_list)
        if self.P != len(ncells):
            raise ValueError("one material entry per problem")
        self.nph = nphase_of(eos_list[0])
        self.nc = ncomp_of(self.nph)
        self.ncells = np.ascontiguousarray(np.asarray(ncells, dtype=np.int64))
        self.offsets = np.concatenate([[0], np.cumsum(self.ncells)])
        self.N = int(self.offsets[-1])
        self.flux = FLUXES[flux] if isinstance(flux, str) else int(flux)
        mats = np.concatenate([materials_array(e) for e in eos_list]).astype(np.float64)
        cf = np.ascontiguousarray(np.broadcast_to(np.asarray(cfl, dtype=np.float64), (self.P,)))
        dxa = None if dx is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(dx, dtype=np.float64), (self.P,)))
        h = ctypes.c_void_p()
        check(lib().hes_batch_create(ctypes.byref(h), self.nph, self.flux, self.P,
                                     self.ncells.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                     ptr(mats), ptr(cf), None if dxa is None else ptr(dxa)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().hes_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self):
        n = ctypes.c_longlong()
        t = ctypes.c_int()
        b = ctypes.c_int()
        x = ctypes.c_int()
        w = ctypes.c_int()
        check(lib().hes_batch_info(self._h, ctypes.byref(n), ctypes.byref(t), ctypes.byref(b),
                                   ctypes.byref(x), ctypes.byref(w)))
        return {"cells": n.value, "tiles": t.value, "block": b.value, "xcd": x.value,
                "tiles_per_wave": w.value}

    def set_state(self, Q, t0=None):
        q, n, _ = columns(Q, self.nc, "Q")
        if n != self.N:
            raise ValueError(f"Q has {n} cells, batch has {self.N}")
        t = None if t0 is None else np.ascontiguousarray(np.broadcast_to(t0, (self.P,)), dtype=np.float64)
        check(lib().hes_batch_set_state(self._h, ptr(q), None if t is None else ptr(t)))

    def set_piecewise(self, states_per_problem, interfaces_per_problem, t0=None):
        """states_per_problem[p]: (ncomp, k_p) conservative states;
        interfaces_per_problem[p]: k_p - 1 sorted positions in (0, 1)."""
        ns = np.array([np.asarray(s).reshape(self.nc, -1).shape[1] for s in states_per_problem],
                      dtype=np.int32)
        st = np.asfortranarray(np.concatenate(
            [np.asarray(s, dtype=np.float64).reshape(self.nc, -1) for s in states_per_problem], axis=1))
        ifc = np.concatenate([np.asarray(f, dtype=np.float64).reshape(-1)
                              for f in interfaces_per_problem] + [np.zeros(1)])
        t = None if t0 is None else np.ascontiguousarray(np.broadcast_to(t0, (self.P,)), dtype=np.float64)
        check(lib().hes_batch_set_piecewise(self._h, ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                            ptr(st), ptr(ifc), None if t is None else ptr(t)))

    def get_state(self):
        out = np.empty((self.nc, self.N), order="F")
        check(lib().hes_batch_get_state(self._h, ptr(out)))
        return out

    def problem_state(self, Q, p):
        return Q[:, self.offsets[p]:self.offsets[p + 1]]

    def times(self):
        t = np.empty(self.P)
        s = np.empty(self.P, dtype=np.int64)
        check(lib().hes_batch_get_times(self._h, ptr(t),
                                        s.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))))
        return t, s

    def set_t_end(self, t_end):
        v = None if t_end is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(t_end, dtype=np.float64), (self.P,)))
        check(lib().hes_batch_set_t_end(self._h, None if v is None else ptr(v)))

    def set_launch(self, block=256, xcd=True, tiles_per_wave=1):
        """Launch shape: block of 64 or 256 threads, XCD-contiguous block order,
        and tiles of 62 cells each wave takes in turn."""
        check(lib().hes_batch_set_launch(self._h, int(block), 1 if xcd else 0, int(tiles_per_wave)))

    def run(self, steps, mode="graph"):
        """Advance every problem by `steps` steps; returns the GPU time in ms."""
        ms = ctypes.c_float()
        check(lib().hes_batch_run(self._h, int(steps), MODES[mode], ctypes.byref(ms)))
        return ms.value

    def autotune(self, steps=3):
        """Time every launch shape (block, XCD order, tiles per wave) on the
        current state, restore the state and keep the fastest."""
        b = ctypes.c_int()
        x = ctypes.c_int()
        w = ctypes.c_int()
        ms = ctypes.c_float()
        check(lib().hes_batch_autotune(self._h, int(steps), ctypes.byref(b), ctypes.byref(x),
                                       ctypes.byref(w), ctypes.byref(ms)))
        return {"block": b.value, "xcd": x.value, "tiles_per_wave": w.value,
                "ms_per_step": ms.value}
