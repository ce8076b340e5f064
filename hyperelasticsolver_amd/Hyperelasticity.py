"""Hyperelasticity: the single-phase module (Hyperelasticity.jl), repaired.

The Julia module keeps a 13-component state
    Q = [rho u (3), rho F (9, row-major: Q[3 i + j] = rho F[i, j], 1-based), rho E]
(Hyperelasticity.jl:70-96) but cannot run as written: it calls `density`,
`energy(S, i)` and `stress(den, e_int, F)` without the equation of state
(Hyperelasticity.jl:29, 91, 104).  These functions take the equation of state
as their first argument, like HyperelasticityMPh, and otherwise keep the
names, layouts and results of the Julia text:

- cons2prim(eos, Q)       -> (den, vel, F (3x3), e_int)   Hyperelasticity.jl:25-36
- prim2cons(eos, vel, F, S) -> Q                           Hyperelasticity.jl:70-96
- flux(eos, Q)            -> physical flux (13)            Hyperelasticity.jl:99-114
- initial_states(eos, testcase) -> (Ql, Qr)                Hyperelasticity.jl:124-177
- postproc_arrays(eos, Q0) -> (den, ent, vel, strs, eint, info)   Hyperelasticity.jl:180-206

The density of a state is rho = sqrt(det(rho F) / rho0), as cons2prim of the
multiphase module computes it (HyperelasticityMPh.jl:113-114).  The library
takes the 13-component layout itself (hes_sp_cons2prim, hes_sp_prim2cons,
hes_sp_flux), so this module only marshals arrays.
"""
import numpy as np

from ._lib import check, columns, lib, ptr
from .EquationsOfState import Barton2009
from .solver import materials_array

def _mat(eos):
    if not isinstance(eos, Barton2009):
        raise NotImplementedError(f"not implemented for EoS: {type(eos).__name__}")
    return materials_array(eos)


def _sp_prims(eos, Q):
    """[den, vel (3), vec(F) (9, column-major), e_int] per 13-component state."""
    mats = _mat(eos)
    q, n, was_vec = columns(Q, 13, "Q")
    P = np.empty((14, n), order="F")
    check(lib().hes_sp_cons2prim(ptr(mats), ptr(q), ptr(P), n))
    return mats, P, n, was_vec


def cons2prim(eos, Q):
    """(den, vel, F, e_int) of a 13-component state (Hyperelasticity.jl:25-36);
    for a (13, n) array: den (n,), vel (3, n), F (3, 3, n), e_int (n,)."""
    _, P, n, was_vec = _sp_prims(eos, Q)
    den, vel, F, e_int = P[0], P[1:4], P[4:13].reshape(3, 3, n, order="F"), P[13]
    if was_vec:
        return float(den[0]), vel[:, 0].copy(), F[:, :, 0].copy(), float(e_int[0])
    return den.copy(), np.asfortranarray(vel), np.asfortranarray(F), e_int.copy()


def prim2cons(eos, vel, F, S):
    """13-component state from velocity, the 3x3 distortion F and entropy S
    (Hyperelasticity.jl:70-96): den = rho0 / det F."""
    mats = _mat(eos)
    P = np.empty(13)
    P[0:3] = np.asarray(vel, dtype=np.float64).reshape(3)
    P[3:12] = np.asarray(F, dtype=np.float64).reshape(3, 3).reshape(-1, order="F")
    P[12] = float(S)
    Q = np.empty(13)
    check(lib().hes_sp_prim2cons(ptr(mats), ptr(P), ptr(Q), 1))
    return Q


def flux(eos, Q):
    """Physical flux of a 13-component state (Hyperelasticity.jl:99-114):
    [rho u1 u - sigma(1,:), 0, 0, 0, rho (F(2,:) u1 - F(1,:) u2),
     rho (F(3,:) u1 - F(1,:) u3), rho u1 E - u . sigma(1,:)]."""
    mats = _mat(eos)
    q, n, was_vec = columns(Q, 13, "Q")
    out = np.empty((13, n), order="F")
    check(lib().hes_sp_flux(ptr(mats), ptr(q), ptr(out), n))
    return out[:, 0].copy() if was_vec else out


# Riemann states of Hyperelasticity.jl:125-172 (u, F rows, S) per side
_S3 = 0.5 * 3 ** 0.5
_CASES = {
    1: ((0.0, 0.5, 1.0), ((0.98, 0.0, 0.0), (0.02, 1.0, 0.1), (0.0, 0.0, 1.0)), 1e-3,
        (0.0, 0.0, 0.0), ((1.0, 0.0, 0.0), (0.0, 1.0, 0.1), (0.0, 0.0, 1.0)), 0.0),
    2: ((2.0, 0.0, 0.1), ((1.0, 0.0, 0.0), (-0.01, 0.95, 0.02), (-0.015, 0.0, 0.9)), 0.0,
        (0.0, -0.03, -0.01), ((1.0, 0.0, 0.0), (0.015, 0.95, 0.0), (-0.01, 0.0, 0.9)), 0.0),
    3: ((1.0, 0.0, 0.0), ((0.5, -_S3, 0.0), (_S3, 0.5, 0.0), (0.0, 0.0, 1.0)), 0.0,
        (1.0, 0.0, 0.0), ((0.5, -_S3, 0.0), (_S3, 0.5, 0.0), (0.0, 0.0, 1.0)), 0.0),
}


def initial_states(eos, testcase):
    """(Ql, Qr) of a test case (Hyperelasticity.jl:124-177); other numbers give
    the state at rest with F = I and S = 0 on both sides."""
    if testcase in _CASES:
        ul, Fl, Sl, ur, Fr, Sr = _CASES[testcase]
    else:
        ul = ur = (0.0, 0.0, 0.0)
        Fl = Fr = np.eye(3)
        Sl = Sr = 0.0
    return prim2cons(eos, ul, Fl, Sl), prim2cons(eos, ur, Fr, Sr)


def postproc_arrays(eos, Q0):
    """Plot arrays of a 13 x nx state (Hyperelasticity.jl:180-206): density,
    entropy, velocity (3, nx), stress (3, 3, nx) with strs[j, k] = sigma[k, j],
    internal energy, and the names tuple."""
    mats, P, n, _ = _sp_prims(eos, Q0)
    den = P[0].copy()
    vel = np.asfortranarray(P[1:4])
    eint = P[13].copy()
    F = np.asfortranarray(P[4:13])
    ent = np.empty(n)
    G = np.empty((9, n), order="F")
    check(lib().hes_finger(ptr(F), ptr(G), n))
    check(lib().hes_entropy(ptr(mats), ptr(np.ascontiguousarray(eint)), ptr(G), ptr(ent), n))
    sig = np.empty((9, n), order="F")
    check(lib().hes_stress_den(ptr(mats), ptr(den), ptr(np.ascontiguousarray(eint)),
                               ptr(F), ptr(sig), n))
    strs = np.transpose(sig.reshape(3, 3, n, order="F"), (1, 0, 2))
    info = ("den", "ent", "vel", "strs", "eint", "info")
    return den, ent, vel, np.asfortranarray(strs), eint, info


__all__ = ["cons2prim", "prim2cons", "flux", "initial_states", "postproc_arrays"]
