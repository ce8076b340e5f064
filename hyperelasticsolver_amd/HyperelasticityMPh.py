"""HyperelasticityMPh: two-phase hyperelasticity (HyperelasticityMPh.jl).

Mirrors the Julia module's functions, names and argument order.  `eos` is a
tuple of two Barton2009 for the mixture functions (`*_mph`, `noncons_flux`,
`get_eigvals(eos::Tuple, ...)`) and a single Barton2009 for the per-phase
methods (`cons2prim`, `flux`, `get_eigvals(eos::T, ...)`), which take one
15-component phase block [a, a rho, a rho u(3), a rho E, a rho vec(A)].
States are 30-vectors or (30, n) arrays; every evaluation runs on the GPU.
"""
import numpy as np

from ._lib import check, columns, lib, ptr
from .EquationsOfState import Barton2009, EoS
from .solver import materials_array


def _pair(eos):
    if not (isinstance(eos, (tuple, list)) and len(eos) == 2):
        raise TypeError("expected a tuple of two equations of state")
    for e in eos:
        if not isinstance(e, Barton2009):
            raise NotImplementedError(f"not implemented for EoS: {type(e).__name__}")
    return materials_array(eos)


def _mph_call(fn, eos, X, in_w, out_w):
    mats = _pair(eos)
    x, n, was_vec = columns(X, in_w)
    out = np.empty((out_w, n), order="F")
    check(fn(2, ptr(mats), ptr(x), ptr(out), n))
    return out[:, 0].copy() if was_vec else out


def prim2cons_mph(eos, P):
    """Primitive -> conservative (HyperelasticityMPh.jl:63-90)."""
    return _mph_call(lib().hes_prim2cons, eos, P, 30, 30)


def cons2prim_mph(eos, Q):
    """Conservative -> primitive [a, rho, u(3), S, vec(A)] x 2 (HyperelasticityMPh.jl:99-133)."""
    return _mph_call(lib().hes_cons2prim, eos, Q, 30, 30)


def flux_mph(eos, Q):
    """Physical flux (HyperelasticityMPh.jl:140-175)."""
    return _mph_call(lib().hes_flux, eos, Q, 30, 30)


def noncons_flux(eos, Q):
    """The 30x30 non-conservative matrix B(Q) (HyperelasticityMPh.jl:178-250);
    (30, 30) for one state, (30, 30, n) for n."""
    mats = _pair(eos)
    x, n, was_vec = columns(Q, 30, "Q")
    out = np.empty((900, n), order="F")
    check(lib().hes_noncons_flux(ptr(mats), ptr(x), ptr(out), n))
    out = out.reshape(30, 30, n, order="F")
    return out[:, :, 0].copy() if was_vec else out


def _phase_call(fn, eos, Q, out_w, *extra):
    if not isinstance(eos, Barton2009):
        raise NotImplementedError(f"not implemented for EoS: {type(eos).__name__}")
    x, n, was_vec = columns(Q, 15, "Q")
    out = np.empty((out_w, n), order="F")
    check(fn(ptr(eos.params()), ptr(x), *extra, ptr(out), n))
    return out[:, 0].copy() if was_vec else out


def cons2prim(eos, Q):
    """Per-phase cons -> prim of a 15-component block (HyperelasticityMPh.jl:106-133):
    [a, rho, u(3), S, vec(A)]."""
    return _phase_call(lib().hes_phase_cons2prim, eos, Q, 15)


def flux(eos, Q):
    """Per-phase physical flux of a 15-component block (HyperelasticityMPh.jl:146-175)."""
    return _phase_call(lib().hes_phase_flux, eos, Q, 15)


def get_eigvals(eos, Q, n):
    """Characteristic speeds along direction n (HyperelasticityMPh.jl:252-266):
    [u.n + c_k, u.n - c_k] per phase (6 per phase), c_k = sqrt|eig(acoustic)|."""
    nv = np.ascontiguousarray(np.asarray(n, dtype=np.float64).reshape(3))
    if isinstance(eos, (tuple, list)):
        mats = _pair(eos)
        x, m, was_vec = columns(Q, 30, "Q")
        out = np.empty((12, m), order="F")
        check(lib().hes_get_eigvals(2, ptr(mats), ptr(x), ptr(nv), ptr(out), m))
        return out[:, 0].copy() if was_vec else out
    return _phase_call(lib().hes_phase_get_eigvals, eos, Q, 6, ptr(nv))


# Riemann test cases of HyperelasticityMPh.jl:275-426: (a_l1, a_l2, a_r1, a_r2, den,
# u_l, F_l (rows), S_l, u_r, F_r (rows), S_r); both phases share u, F and S.
_F_ID = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))
_F_A = ((0.98, 0.0, 0.0), (0.02, 1.0, 0.1), (0.0, 0.0, 1.0))
_F_B = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.1), (0.0, 0.0, 1.0))
_F_C = ((1.0, 0.0, 0.0), (-0.01, 0.95, 0.02), (-0.015, 0.0, 0.9))
_F_D = ((1.0, 0.0, 0.0), (0.015, 0.95, 0.0), (-0.01, 0.0, 0.9))
_CASES = {
    1: (0.5, 0.5, 0.5, 0.5, 5.0, (0, 0, 0), _F_ID, 0.0, (0, 0, 0), _F_ID, 0.0),
    2: (0.5, 0.5, 0.5, 0.5, 5.0, (1.0, 0, 0), _F_ID, 0.0, (1.0, 0, 0), _F_ID, 0.0),
    3: (0.5, 0.5, 0.5, 0.5, 8.9, (2.0, 0.0, 0.1), _F_C, 0.0, (2.0, 0.0, 0.1), _F_C, 0.0),
    4: (0.5, 0.5, 0.5, 0.5, 8.9, (0.0, 0.5, 1.0), _F_A, 1e-3, (0.0, 0.0, 0.0), _F_B, 0.0),
    5: (0.5, 0.5, 0.5, 0.5, 8.9, (2.0, 0.0, 0.1), _F_C, 0.0, (0.0, -0.03, -0.01), _F_D, 0.0),
    6: (0.1, 0.9, 0.9, 0.1, 8.9, (0.0, 0.5, 1.0), _F_A, 1.0e-3, (0.0, 0.0, 0.0), _F_B, 0.0),
    7: (0.1, 0.9, 0.9, 0.1, 8.9, (2.0, 0.0, 0.1), _F_C, 0.0, (0.0, -0.03, -0.01), _F_D, 0.0),
    10: (0.4, 0.6, 0.6, 0.4, 8.9, (2.0, 0.0, 0.1), _F_C, 0.0, (2.0, 0.0, 0.1), _F_C, 0.0),
}


def initial_prims(testcase):
    """Left/right primitive 30-vectors of a Riemann test case."""
    if testcase not in _CASES:
        raise ValueError(f"unknown test case {testcase}")
    al1, al2, ar1, ar2, den, ul, Fl, Sl, ur, Fr, Sr = _CASES[testcase]
    Fl = np.array(Fl, dtype=np.float64)
    Fr = np.array(Fr, dtype=np.float64)
    # den_k = den / det(F_k) (HyperelasticityMPh.jl:412-415)
    dl = den / np.linalg.det(Fl)
    dr = den / np.linalg.det(Fr)

    def block(a, d, u, S, F):
        return [a, d, *map(float, u), S, *F.reshape(-1, order="F")]
    Pl = np.array(block(al1, dl, ul, Sl, Fl) + block(al2, dl, ul, Sl, Fl))
    Pr = np.array(block(ar1, dr, ur, Sr, Fr) + block(ar2, dr, ur, Sr, Fr))
    return Pl, Pr


def initial_states(eos, testcase):
    """Left/right conservative states of a Riemann test case (HyperelasticityMPh.jl:275-426)."""
    Pl, Pr = initial_prims(testcase)
    return prim2cons_mph(eos, Pl), prim2cons_mph(eos, Pr)


__all__ = ["prim2cons_mph", "cons2prim_mph", "flux_mph", "noncons_flux", "initial_states",
           "get_eigvals", "cons2prim", "flux", "initial_prims", "EoS"]
