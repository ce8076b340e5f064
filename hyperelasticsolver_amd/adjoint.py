"""Reverse mode of a fixed-length single-phase Rusanov run on the GPU.

The forward model is the solver's time loop (main.jl:199-225, Rusanov flux
NumFluxes.jl:25-60, CFL step main.jl:204-212) started from a piecewise-constant
state whose pieces are given by their primitives (u, A, S; rho = rho0 / det A,
HyperelasticityMPh.jl:63-87).  The loss is the misfit of the final density
and normal velocity,

    L = sum_i (rho_i - rho*_i)^2 + (u1_i - u1*_i)^2,

and `Adjoint.backward()` returns dL/d(primitives of every piece) and
dL/d(b0, c0).  The CFL step dt = cfl dx / max lam is differentiated through
the argmax (the smallest index among ties), the Rusanov speed max(lam_l,
lam_r) through the larger side (the left one on ties).

Every `every`-th state is stored on the device; the backward pass recomputes
the states in between.  The reference has no adjoint; the Julia-side wrapper
is a ChainRulesCore rrule around the same C-ABI.
"""
import ctypes

import numpy as np

from ._lib import check, lib, ptr
from .solver import materials_array, nphase_of

NPRIM = 13  # u(3), vec(A) column-major (9), S


def phase_prims_to_adjoint(P):
    """Solver primitives [rho, u(3), S, vec(A)] (14 x k) -> adjoint
    primitives [u(3), vec(A), S] (13 x k); rho follows from rho0 / det A."""
    P = np.asarray(P, dtype=np.float64)
    if P.ndim == 1:
        P = P[:, None]
    return np.asfortranarray(np.vstack([P[1:4], P[5:14], P[4:5]]))


class Adjoint:
    def __init__(self, eos, ncells, steps, every=50, cfl=0.6, dx=None):
        if nphase_of(eos) != 1:
            raise ValueError("the adjoint covers single-phase runs")
        self.eos = eos
        self.n = int(ncells)
        self.steps = int(steps)
        self.every = int(every)
        self.cfl = float(cfl)
        self.dx = 1.0 / self.n if dx is None else float(dx)
        self.mats = np.ascontiguousarray(materials_array(eos), dtype=np.float64)
        self.nstates = 0
        self.last_forward_ms = None
        self.last_backward_ms = None
        h = ctypes.c_void_p()
        check(lib().hes_adjoint_create(ctypes.byref(h), self.n, ptr(self.mats), self.cfl, self.dx,
                                       self.steps, self.every))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().hes_adjoint_destroy(self._h)
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

    def set_initial(self, prims, interfaces=()):
        """prims: 13 x k (u(3), vec(A), S per piece); interfaces: k-1 sorted
        positions in (0, 1)."""
        p = np.asfortranarray(np.asarray(prims, dtype=np.float64))
        if p.ndim == 1:
            p = p[:, None]
        if p.shape[0] != NPRIM:
            raise ValueError(f"prims must have {NPRIM} rows")
        k = p.shape[1]
        ifc = np.ascontiguousarray(np.asarray(interfaces, dtype=np.float64).ravel())
        if ifc.size != k - 1:
            raise ValueError("need one interface fewer than pieces")
        check(lib().hes_adjoint_set_initial(self._h, k, ptr(p), ptr(ifc) if ifc.size else None))
        self.nstates = k

    def set_target(self, rho, u1):
        r = np.ascontiguousarray(np.asarray(rho, dtype=np.float64).ravel())
        u = np.ascontiguousarray(np.asarray(u1, dtype=np.float64).ravel())
        if r.size != self.n or u.size != self.n:
            raise ValueError("targets need one value per cell")
        check(lib().hes_adjoint_set_target(self._h, ptr(r), ptr(u)))

    def forward(self):
        """Run the forward model (storing checkpoints); returns the loss."""
        loss = ctypes.c_double()
        ms = ctypes.c_float()
        check(lib().hes_adjoint_forward(self._h, ctypes.byref(loss), ctypes.byref(ms)))
        self.last_forward_ms = ms.value
        return loss.value

    def backward(self):
        """Gradients after `forward()`: (dL/dprims 13 x k, dL/d(b0, c0))."""
        g = np.zeros((NPRIM, self.nstates), order="F")
        gm = np.zeros(2)
        ms = ctypes.c_float()
        check(lib().hes_adjoint_backward(self._h, ptr(g), ptr(gm), ctypes.byref(ms)))
        self.last_backward_ms = ms.value
        return g, gm

    def set_sparse(self, on=True):
        """Skip reverse-step tiles whose cotangent is exactly zero (default on;
        the gradients are the same either way)."""
        check(lib().hes_adjoint_set_sparse(self._h, 1 if on else 0))

    def set_launch(self, use_graph=True):
        """Replay each pass as one hipGraph (default) or launch its kernels one by one."""
        check(lib().hes_adjoint_set_launch(self._h, 1 if use_graph else 0))

    def final_state(self):
        Q = np.empty((14, self.n), order="F")
        check(lib().hes_adjoint_get_state(self._h, ptr(Q)))
        return Q

    def profiles(self):
        """(rho, u1) of the final state as the misfit computes them: the
        target profiles of a run made to be matched by another."""
        r = np.empty(self.n)
        u = np.empty(self.n)
        check(lib().hes_adjoint_get_profiles(self._h, ptr(r), ptr(u)))
        return r, u

    def info(self):
        b = ctypes.c_longlong()
        c = ctypes.c_longlong()
        check(lib().hes_adjoint_info(self._h, ctypes.byref(b), ctypes.byref(c)))
        return {"device_bytes": b.value, "checkpoints": c.value}


class BatchAdjoint:
    """Reverse mode of a batch of independent fixed-length single-phase
    Rusanov runs (one launch per step covers every problem).  Problem p has
    its own material, cell count, dx (default 1 / ncells[p]), pieces and
    target; every problem takes `steps` steps.  `backward()` returns, per
    problem, the gradient of its own misfit L_p with respect to its pieces'
    primitives and its (b0, c0)."""

    def __init__(self, eos_list, ncells, steps, every=50, cfl=0.6, dx=None):
        if any(nphase_of(e) != 1 for e in eos_list):
            raise ValueError("the adjoint covers single-phase runs")
        if len(eos_list) != len(ncells):
            raise ValueError("one material per problem")
        self.eos = list(eos_list)
        self.ncells = np.asarray(ncells, dtype=np.int64)
        self.P = len(self.ncells)
        self.offsets = np.concatenate([[0], np.cumsum(self.ncells)])
        self.N = int(self.offsets[-1])
        self.steps = int(steps)
        self.every = int(every)
        self.cfl = float(cfl)
        self.mats = np.ascontiguousarray(np.concatenate([materials_array(e) for e in self.eos]))
        self.dx = None if dx is None else np.ascontiguousarray(np.asarray(dx, dtype=np.float64))
        self.nstates = None
        self.last_forward_ms = None
        self.last_backward_ms = None
        nc = np.ascontiguousarray(self.ncells)
        h = ctypes.c_void_p()
        check(lib().hes_badjoint_create(
            ctypes.byref(h), self.P, nc.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
            ptr(self.mats), self.cfl, ptr(self.dx) if self.dx is not None else None, self.steps,
            self.every))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().hes_badjoint_destroy(self._h)
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

    def set_initial(self, prims_list, interfaces_list):
        """prims_list[p]: 13 x k_p primitives of problem p's pieces;
        interfaces_list[p]: its k_p - 1 sorted interface positions in (0, 1)."""
        ps, ifs, ks = [], [], []
        for P_, I_ in zip(prims_list, interfaces_list):
            p = np.asarray(P_, dtype=np.float64)
            if p.ndim == 1:
                p = p[:, None]
            if p.shape[0] != NPRIM:
                raise ValueError(f"prims must have {NPRIM} rows")
            i = np.asarray(I_, dtype=np.float64).ravel()
            if i.size != p.shape[1] - 1:
                raise ValueError("need one interface fewer than pieces")
            ps.append(p)
            ifs.append(i)
            ks.append(p.shape[1])
        if len(ks) != self.P:
            raise ValueError("one piece list per problem")
        prims = np.asfortranarray(np.concatenate(ps, axis=1))
        ifc = np.ascontiguousarray(np.concatenate(ifs)) if sum(k - 1 for k in ks) else None
        k = np.ascontiguousarray(np.asarray(ks, dtype=np.int32))
        check(lib().hes_badjoint_set_initial(self._h, k.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                             ptr(prims), ptr(ifc) if ifc is not None else None))
        self.nstates = ks

    def set_target(self, rho, u1):
        """Target profiles over the packed cells (N values each)."""
        r = np.ascontiguousarray(np.asarray(rho, dtype=np.float64).ravel())
        u = np.ascontiguousarray(np.asarray(u1, dtype=np.float64).ravel())
        if r.size != self.N or u.size != self.N:
            raise ValueError("targets need one value per packed cell")
        check(lib().hes_badjoint_set_target(self._h, ptr(r), ptr(u)))

    def forward(self):
        """Run every problem forward (storing checkpoints); returns the
        per-problem losses L_p."""
        losses = np.zeros(self.P)
        tot = ctypes.c_double()
        ms = ctypes.c_float()
        check(lib().hes_badjoint_forward(self._h, ptr(losses), ctypes.byref(tot), ctypes.byref(ms)))
        self.last_forward_ms = ms.value
        return losses

    def backward(self):
        """Gradients after `forward()`: a list of (dL_p/dprims 13 x k_p,
        dL_p/d(b0_p, c0_p)) per problem."""
        g = np.zeros((NPRIM, sum(self.nstates)), order="F")
        gm = np.zeros(2 * self.P)
        ms = ctypes.c_float()
        check(lib().hes_badjoint_backward(self._h, ptr(g), ptr(gm), ctypes.byref(ms)))
        self.last_backward_ms = ms.value
        out, c = [], 0
        for p, k in enumerate(self.nstates):
            out.append((g[:, c:c + k].copy(order="F"), gm[2 * p:2 * p + 2].copy()))
            c += k
        return out

    def set_sparse(self, on=True):
        check(lib().hes_badjoint_set_sparse(self._h, 1 if on else 0))

    def set_launch(self, use_graph=True):
        check(lib().hes_badjoint_set_launch(self._h, 1 if use_graph else 0))

    def final_state(self):
        Q = np.empty((14, self.N), order="F")
        check(lib().hes_badjoint_get_state(self._h, ptr(Q)))
        return Q

    def profiles(self):
        """(rho, u1) of the final packed state as the misfit computes them."""
        r = np.empty(self.N)
        u = np.empty(self.N)
        check(lib().hes_badjoint_get_profiles(self._h, ptr(r), ptr(u)))
        return r, u

    def info(self):
        n = ctypes.c_longlong()
        b = ctypes.c_longlong()
        c = ctypes.c_longlong()
        check(lib().hes_badjoint_info(self._h, ctypes.byref(n), ctypes.byref(b), ctypes.byref(c)))
        return {"total_cells": n.value, "device_bytes": b.value, "checkpoints": c.value}
