"""NumFluxes: path-conservative numerical fluxes (NumFluxes.jl).

Same names and argument order as the Julia module: `eos` is a tuple of two
Barton2009, states are 30-vectors (or (30, n) arrays for n faces at once), and
each function returns the triple (cons, noncons_minus, noncons_plus) that the
time update consumes (main.jl:43-60).  Every evaluation runs on the GPU.

- lxf(eos, Q_l, Q_r, lambda) -- NumFluxes.jl:25-60: the Rusanov flux
  1/2 (F_l + F_r) - lambda/2 (Q_r - Q_l) and D- = D+ = 1/2 of the Gauss-Lobatto-6
  integral of B(psi) dpsi along the straight path psi(s) = Q_l (1 - s) + Q_r s.
- hll(eos, Q_l, Q_r, eigvals) -- NumFluxes.jl:70-132: D- and D+ of the
  path-conservative HLL scheme (Gauss-Legendre-6 path integrals); `eigvals` is
  the pair (eigenvalues of the left cell, eigenvalues of the right cell) as
  main.jl passes them, and the conservative part is zero.

The one-phase forms (eos a single Barton2009, 14-component states) are the same
fluxes with the path integrals absent.
"""
import numpy as np

from ._lib import FLUX_HLL, FLUX_RUSANOV, check, columns, lib, ptr
from .EquationsOfState import Barton2009
from .solver import materials_array


def _setup(eos, Q_l, Q_r):
    if isinstance(eos, (tuple, list)):
        if len(eos) != 2 or not all(isinstance(e, Barton2009) for e in eos):
            raise NotImplementedError("expected a tuple of two Barton2009")
        nph, nc = 2, 30
    elif isinstance(eos, Barton2009):
        nph, nc = 1, 14
    else:
        raise NotImplementedError(f"not implemented for EoS: {type(eos).__name__}")
    ql, n, was_vec = columns(Q_l, nc, "Q_l")
    qr, n2, _ = columns(Q_r, nc, "Q_r")
    if n != n2:
        raise ValueError("Q_l and Q_r hold different numbers of faces")
    return nph, nc, materials_array(eos), ql, qr, n, was_vec


def _call(kind, eos, Q_l, Q_r, aux_fn):
    nph, nc, mats, ql, qr, n, was_vec = _setup(eos, Q_l, Q_r)
    aux = aux_fn(n)
    cons = np.empty((nc, n), order="F")
    dm = np.empty((nc, n), order="F")
    dp = np.empty((nc, n), order="F")
    check(lib().hes_face_flux(nph, kind, ptr(mats), ptr(ql), ptr(qr), ptr(aux), ptr(cons), ptr(dm),
                              ptr(dp), n))
    if was_vec:
        return cons[:, 0].copy(), dm[:, 0].copy(), dp[:, 0].copy()
    return cons, dm, dp


def lxf(eos, Q_l, Q_r, lam):
    """Rusanov / local Lax-Friedrichs face (NumFluxes.jl:25-60); lam is the
    face speed (a scalar, or one per face)."""
    def aux(n):
        return np.ascontiguousarray(np.broadcast_to(np.asarray(lam, dtype=np.float64), (n,)))
    return _call(FLUX_RUSANOV, eos, Q_l, Q_r, aux)


def hll(eos, Q_l, Q_r, eigvals):
    """Path-conservative HLL face (NumFluxes.jl:70-132).  eigvals = (eigenvalues
    of the left state, eigenvalues of the right state); for n faces each may be
    a (k, n) array.  Only minimum(eigvals[0]) and maximum(eigvals[1]) enter
    (NumFluxes.jl:90-91)."""
    ev_l = np.asarray(eigvals[0], dtype=np.float64)
    ev_r = np.asarray(eigvals[1], dtype=np.float64)

    def aux(n):
        lo = np.broadcast_to(ev_l.min(axis=0) if ev_l.ndim > 1 else ev_l.min(), (n,))
        hi = np.broadcast_to(ev_r.max(axis=0) if ev_r.ndim > 1 else ev_r.max(), (n,))
        return np.ascontiguousarray(np.stack([lo, hi], axis=0).T.ravel())
    return _call(FLUX_HLL, eos, Q_l, Q_r, aux)


__all__ = ["lxf", "hll"]
