"""EquationsOfState: the Barton2009 equation of state (EquationsOfState.jl).

Same names, keyword arguments and argument order as the Julia module.  Tensors
are column-major 9-vectors (Julia `vec`), or (9, n) arrays for n points at once;
every evaluation runs on the GPU through libhes.so.
"""
import numpy as np

from . import _lib
from ._lib import check, columns, lib, ptr


class EoS:
    """Abstract equation of state (EquationsOfState.jl:26)."""

    def params(self):
        raise NotImplementedError(f"params() isn't implemented for EoS: {type(self).__name__}")


class Barton2009(EoS):
    """Barton2009 EoS (EquationsOfState.jl:71-116); keyword defaults of :90."""

    def __init__(self, *, _rho0=8.93, _c0=4.6, _cv=3.9e-4, _t0=300, _b0=2.1, _alpha=1, _beta=3,
                 _gamma=2):
        self.rho0 = _rho0
        self.c0 = _c0
        self.cv = _cv
        self.t0 = _t0
        self.b0 = _b0
        self.alpha = _alpha
        self.beta = _beta
        self.gamma = _gamma
        # secondary parameters (EquationsOfState.jl:111-112)
        self.b0sq = _b0 ** 2
        self.k0 = _c0 ** 2 - (4 / 3) * _b0 ** 2

    def params(self):
        """The 8 primary constants as the C-ABI takes them."""
        return np.array([self.rho0, self.c0, self.cv, self.t0, self.b0, self.alpha, self.beta,
                         self.gamma], dtype=np.float64)

    def __repr__(self):
        return (f"Barton2009(_rho0={self.rho0}, _c0={self.c0}, _cv={self.cv}, _t0={self.t0}, "
                f"_b0={self.b0}, _alpha={self.alpha}, _beta={self.beta}, _gamma={self.gamma})")

    def __eq__(self, other):
        return isinstance(other, Barton2009) and np.array_equal(self.params(), other.params())

    def __hash__(self):
        return hash(tuple(self.params()))


def _barton(eos, fname):
    if not isinstance(eos, Barton2009):
        raise NotImplementedError(f"{fname}() isn't implemented for EoS: {type(eos).__name__}")
    return eos.params()


def _scalar_or(out, was_vec):
    return float(out[0]) if was_vec else out


def _vals(x, n):
    v = np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (n,)))
    return v


def energy(eos, S, G):
    """Internal energy e(S, G) (EquationsOfState.jl:118-137)."""
    mat = _barton(eos, "energy")
    g, n, was_vec = columns(G, 9, "G")
    s = _vals(S, n)
    out = np.empty(n)
    check(lib().hes_energy(ptr(mat), ptr(s), ptr(g), ptr(out), n))
    return _scalar_or(out, was_vec and np.ndim(S) == 0)


def entropy(eos, e_int, G):
    """Entropy S(e_int, G) with the 1e-6 clamp (EquationsOfState.jl:139-156)."""
    mat = _barton(eos, "entropy")
    g, n, was_vec = columns(G, 9, "G")
    e = _vals(e_int, n)
    out = np.empty(n)
    check(lib().hes_entropy(ptr(mat), ptr(e), ptr(g), ptr(out), n))
    return _scalar_or(out, was_vec and np.ndim(e_int) == 0)


def temperature(eos, S, G):
    """de/dS at fixed G (HyperelasticityMPh.jl:212)."""
    mat = _barton(eos, "temperature")
    g, n, was_vec = columns(G, 9, "G")
    s = _vals(S, n)
    out = np.empty(n)
    check(lib().hes_temperature(ptr(mat), ptr(s), ptr(g), ptr(out), n))
    return _scalar_or(out, was_vec and np.ndim(S) == 0)


def stress(eos, *args):
    """Cauchy stress as a column-major 9-vector.

    stress(eos, ent, F)        -- EquationsOfState.jl:179-190 (den = rho0/det F)
    stress(eos, den, e_int, F) -- EquationsOfState.jl:162-176
    """
    mat = _barton(eos, "stress")
    if len(args) == 2:
        ent, F = args
        f, n, was_vec = columns(F, 9, "F")
        s = _vals(ent, n)
        out = np.empty((9, n), order="F")
        check(lib().hes_stress(ptr(mat), ptr(s), ptr(f), ptr(out), n))
    elif len(args) == 3:
        den, e_int, F = args
        f, n, was_vec = columns(F, 9, "F")
        d = _vals(den, n)
        e = _vals(e_int, n)
        out = np.empty((9, n), order="F")
        check(lib().hes_stress_den(ptr(mat), ptr(d), ptr(e), ptr(f), ptr(out), n))
    else:
        raise TypeError("stress(eos, ent, F) or stress(eos, den, e_int, F)")
    return out[:, 0].copy() if was_vec else out


def acoustic(eos, *args):
    """Acoustic tensor (3x3, or (3, 3, n)).

    acoustic(eos, ent, F, n) -- EquationsOfState.jl:223-246
    acoustic(eos, P, n)      -- EquationsOfState.jl:192-221 (P: 15 phase primitives)
    """
    mat = _barton(eos, "acoustic")
    if len(args) == 3:
        ent, F, nv = args
        f, n, was_vec = columns(F, 9, "F")
        s = _vals(ent, n)
        nn = np.asfortranarray(np.broadcast_to(np.asarray(nv, dtype=np.float64).reshape(3, -1), (3, n)))
        out = np.empty((9, n), order="F")
        check(lib().hes_acoustic(ptr(mat), ptr(s), ptr(f), ptr(nn), ptr(out), n))
    elif len(args) == 2:
        P, nv = args
        p, n, was_vec = columns(P, 15, "P")
        nn = np.asfortranarray(np.broadcast_to(np.asarray(nv, dtype=np.float64).reshape(3, -1), (3, n)))
        out = np.empty((9, n), order="F")
        check(lib().hes_acoustic_prim(ptr(mat), ptr(p), ptr(nn), ptr(out), n))
    else:
        raise TypeError("acoustic(eos, ent, F, n) or acoustic(eos, P, n)")
    out = out.reshape(3, 3, n, order="F")
    return out[:, :, 0].copy() if was_vec else out


__all__ = ["EoS", "Barton2009", "energy", "entropy", "stress", "acoustic", "temperature"]
